// m2d_adjoint.hip -- reverse-mode gradient of the self-consistent 2D step
// (Pic::advance with a solve every step, pic.cpp:330-358, one species,
// optionally with MCC collisions against continuum targets) with respect to
// the voltages of groups of FIXED electrode nodes, for the objective J =
// mean kinetic energy in eV of the live particles after the last step
// (energy_dist.mean_tot, histogram.cpp:68-71).
//
// Forward step n (the library's own kernels):
//   u^n = A^-1 b(rho^n, V)            Fields::boundary_solve (fields.cpp:278-312)
//   a = (q dt / 2m) E(x^n; u^n)       E = -(staggered gradient gather)
//   v^{n+1/2} = M (v^{n-1/2} + a) + a Boris with the constant rotation M
//   x^{n+1} = x^n + v^{n+1/2} dt      (x and z)
//   v^{n+1/2} <- S(v^{n+1/2}) if the slot's MCC trial fired (after the drift)
//   boundary; rho^{n+1} = sum q W(x^{n+1})  CIC
// Collisions are differentiated under common random numbers: the trial
// word, the target's velocity and the scattering angles are functions of
// (seed, stream, slot, step) only, so a run at other voltages draws the same
// numbers; which collisions fire does not depend on the voltages (the trial
// probability is fixed by the lifetime), and the interaction picked is taken
// as fixed (it is a step function of the relative speed, whose derivative
// is zero almost everywhere).  S's Jacobian comes from the device scatter
// itself evaluated on forward-mode dual numbers.
// Reverse step n, for the particles alive after the step (removals are
// frozen):
//   lx += q grad(interp lrho^{n+1})(x^{n+1})
//   lv <- S'(v)^T lv for a slot whose trial fired (v: the velocity before
//   the scatter, recomputed in FP64 from v^{n-1/2} and u^n);  lv += dt lx
//   la = lv + M^T lv;  lv <- M^T lv
//   lu += -(q dt/2m) G(x^n)^T la;  lx += -(q dt/2m) (dE/dx)^T la
//   A_FF y = lu_F (same multigrid solver, zero Dirichlet values)
//   lrho^n = -dx^2/(eps0 dV) y on free nodes;
//   dJ/dV_D += lu_D - sum over free neighbours y   (u_F depends on u_D
//   through -A_FD u_D; A_FF is symmetric, so the adjoint solve is a solve)
// The forward stores u^n every step and the particle store every K steps;
// the reverse sweep replays the pushes of one K-step segment at a time from
// its checkpoint (bit for bit: same kernels, same u^n) to get x^n.
#include "m2d_scatter.hpp"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <new>
#include <vector>

namespace m2d {

struct AdjArgs {
    const float *x0, *z0;        // positions before the step (x^n)
    const float *x1, *z1;        // after the step (x^{n+1}); dead bits mark removals
    int64_t n;
    double *lx, *lz, *lvx, *lvy, *lvz;
    const double *lrho1;         // adjoint of rho^{n+1}
    const double *u;             // u^n (FP64)
    double *lu;                  // adjoint of u^n (accumulated)
    int nx, nz;
    double idx, idz, xmin, zmin;
    double q, dt, c;             // charge, dt, q dt / 2m
    double M[9];                 // Boris rotation, row-major: v' = M v
    unsigned long long *counters;
    // MCC (mcc non-null): v^{n-1/2} before the step and the draw keys
    const float *v0x, *v0y, *v0z;
    const McDev *mcc;
    uint64_t thr, seed, step;    // trial threshold; Philox key; niter of this step
    uint32_t stream;
};

__device__ __forceinline__ double uat(const double *u, int nz, int i, int j) { return u[(size_t)i * nz + j]; }

// lu contributions of one particle: 8 node updates of the x component and 8
// of the z component of the gather transpose
struct LuAdd {
    int row[16], col[16];
    double val[16];
    int n = 0;
    __device__ void add(int r, int c, double v) { row[n] = r; col[n] = c; val[n] = v; n++; }
};

constexpr int kAdjWin = 2304;  // doubles of the block's LDS window (18 KiB)

// Fields::E at x^n in FP64 from u^n through the staggered edge gradients,
// with the pieces its adjoint needs.  Returns true when the point lies
// outside the grid (it is clamped onto it, as the forward gather does).
struct AdjGather {
    int jx, i0, er[2], ec[2];
    double fx, fyx, fxz, fz;
    double g[2][2], h[2][2];  // x and z edge gradients
    double Ex, Ez;
};

__device__ __forceinline__ bool adj_gather(const AdjArgs &a, float xa, float za, AdjGather &G)
{
    const int nx = a.nx, nz = a.nz;
    double X = ((double)xa - a.xmin) * a.idx, Z = ((double)za - a.zmin) * a.idz;
    const bool bad = !(X > -1.0 && X < nx - 0.5 && Z > -1.0 && Z < nz - 0.5);
    if (bad) {
        X = fmin(fmax(X, 0.0), (double)(nx - 1));
        Z = fmin(fmax(Z, 0.0), (double)(nz - 1));
    }
    const int ih = max(0, min((int)(X + 0.5), nx - 1)), jx = max(0, min((int)Z, nz - 2));
    const int i0 = max(0, min((int)X, nx - 2)), jh = max(0, min((int)(Z + 0.5), nz - 1));
    G.jx = jx;
    G.i0 = i0;
    G.fx = X - ih + 0.5; G.fyx = Z - jx; G.fxz = X - i0; G.fz = Z - jh + 0.5;
    // x component: edge e between rows ee-1 and ee, ee = clamp(e, 1, nx-1)
    for (int c = 0; c < 2; c++) G.er[c] = min(max(ih + c, 1), nx - 1);
    for (int c = 0; c < 2; c++)
        for (int d = 0; d < 2; d++)
            G.g[c][d] = (uat(a.u, nz, G.er[c], jx + d) - uat(a.u, nz, G.er[c] - 1, jx + d)) * a.idx;
    // z component: edge e between columns ee-1 and ee, ee = clamp(e, 1, nz-1)
    for (int d = 0; d < 2; d++) G.ec[d] = min(max(jh + d, 1), nz - 1);
    for (int c = 0; c < 2; c++)
        for (int d = 0; d < 2; d++)
            G.h[c][d] = (uat(a.u, nz, i0 + c, G.ec[d]) - uat(a.u, nz, i0 + c, G.ec[d] - 1)) * a.idz;
    const double fx = G.fx, fyx = G.fyx, fxz = G.fxz, fz = G.fz;
    G.Ex = -((1 - fx) * ((1 - fyx) * G.g[0][0] + fyx * G.g[0][1]) + fx * ((1 - fyx) * G.g[1][0] + fyx * G.g[1][1]));
    G.Ez = -((1 - fxz) * ((1 - fz) * G.h[0][0] + fz * G.h[0][1]) + fxz * ((1 - fz) * G.h[1][0] + fz * G.h[1][1]));
    return bad;
}

// Adjoint of the MCC scatter of step n, before k_adj_step2d of the same
// step: for each slot alive after the step whose trial fired, lv <- S'^T lv
// with S' the Jacobian of BaseSpecies::scatter at the velocity it scattered
// (the Boris push of v^{n-1/2} with E(x^n; u^n), in FP64) for the slot's own
// draws.
__global__ void __launch_bounds__(256) k_adj_scatter2d(AdjArgs a)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    const float xa = a.x0[p];
    if (is_dead(xa) || is_dead(a.x1[p])) return;  // not alive, or removed by the step: no scatter
    if ((uint64_t)trial_word(a.seed, a.stream, (uint64_t)p, a.step) >= a.thr) return;
    AdjGather G;
    (void)adj_gather(a, xa, a.z0[p], G);
    // v = M (v^{n-1/2} + c E) + c E
    const double ax = a.c * G.Ex, az = a.c * G.Ez;
    const double w0 = (double)a.v0x[p] + ax, w1 = (double)a.v0y[p], w2 = (double)a.v0z[p] + az;
    const double *M = a.M;
    Dual3 vx = mk(M[0] * w0 + M[1] * w1 + M[2] * w2 + ax, 1.0, 0.0, 0.0);
    Dual3 vy = mk(M[3] * w0 + M[4] * w1 + M[5] * w2, 0.0, 1.0, 0.0);
    Dual3 vz = mk(M[6] * w0 + M[7] * w1 + M[8] * w2 + az, 0.0, 0.0, 1.0);
    Rng rg;
    rg.begin(a.seed, a.stream, (uint32_t)p, a.step, kPurposeMcc);
    scatter_d(*a.mcc, rg, vx, vy, vz, a.counters);
    const double l0 = a.lvx[p], l1 = a.lvy[p], l2 = a.lvz[p];
    a.lvx[p] = l0 * vx.d[0] + l1 * vy.d[0] + l2 * vz.d[0];
    a.lvy[p] = l0 * vx.d[1] + l1 * vy.d[1] + l2 * vz.d[1];
    a.lvz[p] = l0 * vx.d[2] + l1 * vy.d[2] + l2 * vz.d[2];
}

// one reverse step for every particle (see the header comment); the gather
// transpose is summed in an LDS window over the block's node bounding box
// when it fits, else added to lu directly
__global__ void __launch_bounds__(256) k_adj_step2d(AdjArgs a)
{
    __shared__ double win[kAdjWin];
    __shared__ int bb[4];  // row min, row max, col min, col max
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nx = a.nx, nz = a.nz;
    LuAdd L;
    bool active = false;
    if (threadIdx.x == 0) { bb[0] = nx; bb[1] = -1; bb[2] = nz; bb[3] = -1; }
    if (p < a.n) {
        const float xa = a.x0[p];
        const float xb = a.x1[p];
        if (!is_dead(xa) && is_dead(xb)) {
            // removed by this step: nothing after it depends on it
            a.lx[p] = a.lz[p] = a.lvx[p] = a.lvy[p] = a.lvz[p] = 0.0;
        } else if (!is_dead(xa)) {
            active = true;
            double lx = a.lx[p], lz = a.lz[p], lvx = a.lvx[p], lvy = a.lvy[p], lvz = a.lvz[p];
            // ---- deposit adjoint at x^{n+1}: gradient of the CIC interpolation of lrho
            {
                const double X = ((double)xb - a.xmin) * a.idx, Z = ((double)a.z1[p] - a.zmin) * a.idz;
                if (X >= 0.0 && X < nx - 1.0 && Z >= 0.0 && Z < nz - 1.0) {
                    const int i = (int)X, j = (int)Z;
                    const double u = X - i, v = Z - j;
                    const double l00 = uat(a.lrho1, nz, i, j), l10 = uat(a.lrho1, nz, i + 1, j);
                    const double l01 = uat(a.lrho1, nz, i, j + 1), l11 = uat(a.lrho1, nz, i + 1, j + 1);
                    lx += a.q * a.idx * ((1 - v) * (l10 - l00) + v * (l11 - l01));
                    lz += a.q * a.idz * ((1 - u) * (l01 - l00) + u * (l11 - l10));
                }
            }
            // ---- drift x^{n+1} = x^n + v dt (x, z)
            lvx += a.dt * lx;
            lvz += a.dt * lz;
            // ---- Boris: v' = M (v + a) + a, a = (ax, 0, az)
            const double *M = a.M;
            const double mtx = M[0] * lvx + M[3] * lvy + M[6] * lvz;
            const double mty = M[1] * lvx + M[4] * lvy + M[7] * lvz;
            const double mtz = M[2] * lvx + M[5] * lvy + M[8] * lvz;
            const double lax = lvx + mtx, laz = lvz + mtz;
            lvx = mtx; lvy = mty; lvz = mtz;
            // ---- gather adjoint at x^n: a = c E, E = -(staggered gradient of u)
            const double lEx = a.c * lax, lEz = a.c * laz;
            AdjGather G;
            if (adj_gather(a, xa, a.z0[p], G)) raise_flag(a.counters, M2D_FLAG_OUT_OF_RANGE);
            const int jx = G.jx, i0 = G.i0;
            const int *er = G.er, *ec = G.ec;
            const double fx = G.fx, fyx = G.fyx, fxz = G.fxz, fz = G.fz;
            const double(&g)[2][2] = G.g;
            const double(&h)[2][2] = G.h;
            // E_x = -sum wx_cd g_cd, E_z = -sum wz_cd h_cd
            const double wx[2][2] = {{(1 - fx) * (1 - fyx), (1 - fx) * fyx}, {fx * (1 - fyx), fx * fyx}};
            const double wz[2][2] = {{(1 - fxz) * (1 - fz), (1 - fxz) * fz}, {fxz * (1 - fz), fxz * fz}};
            // position derivatives of E (one-sided at cell faces)
            const double dExdx =
                -a.idx * (-((1 - fyx) * g[0][0] + fyx * g[0][1]) + ((1 - fyx) * g[1][0] + fyx * g[1][1]));
            const double dExdz = -a.idz * ((1 - fx) * (g[0][1] - g[0][0]) + fx * (g[1][1] - g[1][0]));
            const double dEzdx =
                -a.idx * (-((1 - fz) * h[0][0] + fz * h[0][1]) + ((1 - fz) * h[1][0] + fz * h[1][1]));
            const double dEzdz = -a.idz * ((1 - fxz) * (h[0][1] - h[0][0]) + fxz * (h[1][1] - h[1][0]));
            lx += lEx * dExdx + lEz * dEzdx;
            lz += lEx * dExdz + lEz * dEzdz;
            // transpose of the gather
            for (int c = 0; c < 2; c++)
                for (int d = 0; d < 2; d++) {
                    const double tx = -lEx * wx[c][d] * a.idx;
                    L.add(er[c], jx + d, tx);
                    L.add(er[c] - 1, jx + d, -tx);
                    const double tz = -lEz * wz[c][d] * a.idz;
                    L.add(i0 + c, ec[d], tz);
                    L.add(i0 + c, ec[d] - 1, -tz);
                }
            a.lx[p] = lx; a.lz[p] = lz;
            a.lvx[p] = lvx; a.lvy[p] = lvy; a.lvz[p] = lvz;
        }
    }
    __syncthreads();
    if (active) {
        int r0 = nx, r1 = -1, c0 = nz, c1 = -1;
        for (int k = 0; k < L.n; k++) {
            r0 = min(r0, L.row[k]); r1 = max(r1, L.row[k]);
            c0 = min(c0, L.col[k]); c1 = max(c1, L.col[k]);
        }
        atomicMin(&bb[0], r0); atomicMax(&bb[1], r1);
        atomicMin(&bb[2], c0); atomicMax(&bb[3], c1);
    }
    __syncthreads();
    const int R0 = bb[0], C0 = bb[2], W = bb[3] - bb[2] + 1, H = bb[1] - bb[0] + 1;
    if (H <= 0) return;  // uniform: no active particle in the block
    if (H * W <= kAdjWin) {
        for (int k = threadIdx.x; k < H * W; k += blockDim.x) win[k] = 0.0;
        __syncthreads();
        for (int k = 0; k < L.n; k++)
            if (L.val[k] != 0.0) atomicAdd(&win[(L.row[k] - R0) * W + (L.col[k] - C0)], L.val[k]);
        __syncthreads();
        for (int k = threadIdx.x; k < H * W; k += blockDim.x) {
            const double v = win[k];
            if (v != 0.0) atomicAdd(a.lu + (size_t)(R0 + k / W) * nz + (C0 + k % W), v);
        }
    } else {
        for (int k = 0; k < L.n; k++)
            if (L.val[k] != 0.0) atomicAdd(a.lu + (size_t)L.row[k] * nz + L.col[k], L.val[k]);
    }
}

// final condition: lv = dJ/dv = m v / (q_e N_live) on live particles
__global__ void k_adj_init2d(const float *x, const float *vx, const float *vy, const float *vz, int64_t n,
                             double scale, double *lx, double *lz, double *lvx, double *lvy, double *lvz)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const bool live = !is_dead(x[p]);
    lx[p] = 0.0;
    lz[p] = 0.0;
    lvx[p] = live ? scale * vx[p] : 0.0;
    lvy[p] = live ? scale * vy[p] : 0.0;
    lvz[p] = live ? scale * vz[p] : 0.0;
}

__device__ __forceinline__ bool dir2(uint8_t m) { return m == M2D_FIXED || m == M2D_FIXED_RF; }

// rhs of the adjoint solve: lu on free rows, 0 on Dirichlet rows
__global__ void k_adj_rhs(const uint8_t *mask, const double *lu, double *rhs, size_t n)
{
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) rhs[k] = dir2(mask[k]) ? 0.0 : lu[k];
}

// lrho^n = scale * y on free nodes
__global__ void k_adj_lrho(const uint8_t *mask, const double *y, double scale, double *lrho, size_t n)
{
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) lrho[k] = dir2(mask[k]) ? 0.0 : scale * y[k];
}

constexpr int kMaxGroups = 64;

// per-block sums over the nodes of each group of lu_D - sum_free_nb y
__global__ void __launch_bounds__(256) k_adj_groups(const uint8_t *mask, const int *group, const double *lu,
                                                    const double *y, int nx, int nz, int n_groups, double *partial)
{
    __shared__ double s[kMaxGroups];
    for (int k = threadIdx.x; k < n_groups; k += blockDim.x) s[k] = 0.0;
    __syncthreads();
    const size_t n = (size_t)nx * nz;
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) {
        const int gk = group[k];
        if (gk >= 0) {
            const int i = (int)(k / nz), j = (int)(k % nz);
            double v = lu[k];
            if (i > 0 && !dir2(mask[k - nz])) v -= y[k - nz];
            if (i < nx - 1 && !dir2(mask[k + nz])) v -= y[k + nz];
            if (j > 0 && !dir2(mask[k - 1])) v -= y[k - 1];
            if (j < nz - 1 && !dir2(mask[k + 1])) v -= y[k + 1];
            atomicAdd(&s[gk], v);
        }
    }
    __syncthreads();
    for (int g = threadIdx.x; g < n_groups; g += blockDim.x) partial[(size_t)blockIdx.x * n_groups + g] = s[g];
}

// grad[g] += the sum over blocks of group g's partials, one workgroup per
// group: thread t sums blocks t, t + 256, ... in order, then a fixed tree
// over the threads (the same order every run)
__global__ void __launch_bounds__(256) k_adj_groups_sum(const double *partial, int nb, int n_groups, double *grad)
{
    __shared__ double s[256];
    const int g = blockIdx.x;
    double v = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) v += partial[(size_t)b * n_groups + g];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) grad[g] += s[0];
}

// device buffer freed at scope exit
struct DBuf {
    void *p = nullptr;
    ~DBuf()
    {
        if (p) {
            (void)hipDeviceSynchronize();  // no queued kernel may still use it
            (void)hipFree(p);
        }
    }
    int alloc(size_t bytes)
    {
        hipError_t e = hipMalloc(&p, bytes < 16 ? 16 : bytes);
        if (e != hipSuccess) {
            p = nullptr;
            return set_error(M2D_ERR_NO_MEMORY, "adjoint: hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
        }
        return M2D_OK;
    }
    template <typename T> T *as() const { return (T *)p; }
};

static unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace m2d

using namespace m2d;

extern "C" {

int m2d_adjoint_energy_2d(m2d_ctx *ctx, m2d_fields2d *f, const m2d_species2d *s, const m2d_particles2d *p,
                          const m2d_adjoint_cfg *cfg, double *J, double *grad, m2d_adjoint_info *info)
{
    using clock = std::chrono::steady_clock;
    M2D_CHECK(ctx && f && s && p && cfg && J, M2D_ERR_BAD_ARG, "m2d_adjoint_energy_2d: NULL argument");
    M2D_CHECK(f->ctx == ctx, M2D_ERR_BAD_ARG, "m2d_adjoint_energy_2d: fields belong to another context");
    M2D_CHECK(!f->bt, M2D_ERR_UNSUPPORTED, "m2d_adjoint_energy_2d: gridded B is not supported");
    M2D_CHECK(!p->xl, M2D_ERR_UNSUPPORTED, "m2d_adjoint_energy_2d: long-time positions are not supported");
    M2D_CHECK(f->p.selfconsistent && !f->p.rf && !f->p.electric_field_from_file, M2D_ERR_UNSUPPORTED,
              "m2d_adjoint_energy_2d: needs a self-consistent field without RF and without a field file");
    M2D_CHECK(cfg->steps >= 1 && cfg->checkpoint_every >= 1, M2D_ERR_BAD_ARG,
              "m2d_adjoint_energy_2d: steps and checkpoint_every must be >= 1");
    M2D_CHECK(!grad || (cfg->n_groups >= 1 && cfg->n_groups <= kMaxGroups && cfg->group), M2D_ERR_BAD_ARG,
              "m2d_adjoint_energy_2d: 1..%d electrode groups with a node map", kMaxGroups);
    M2D_CHECK(p->n > 0 && p->x && p->z && p->vx && p->vy && p->vz, M2D_ERR_BAD_ARG,
              "m2d_adjoint_energy_2d: empty particle store");
    M2D_CHECK(cfg->adjoint_tol >= 0 && cfg->adjoint_tol < 1, M2D_ERR_BAD_ARG,
              "m2d_adjoint_energy_2d: adjoint_tol %g out of [0, 1)", cfg->adjoint_tol);
    m2d_mcc *mcc = cfg->mcc;
    if (mcc) {
        M2D_CHECK(mcc->ctx == ctx, M2D_ERR_BAD_ARG, "m2d_adjoint_energy_2d: MCC tables belong to another context");
        M2D_CHECK(s->lifetime > 0 && std::isfinite(s->lifetime), M2D_ERR_BAD_ARG,
                  "m2d_adjoint_energy_2d: MCC needs the species' finite lifetime");
        for (int t = 0; t < mcc->h.n_targets; t++)
            M2D_CHECK(mcc->h.t[t].px == nullptr, M2D_ERR_UNSUPPORTED,
                      "m2d_adjoint_energy_2d: targets with a live partner ensemble are not differentiated");
        M2D_TRY(mcc_sync(mcc));
    }
    const double adj_tol = cfg->adjoint_tol > 0 ? cfg->adjoint_tol : 1e-10;
    const int N = cfg->steps, K = cfg->checkpoint_every;
    const size_t nc = f->n, np = (size_t)p->n;
    hipStream_t st = ctx->stream;
    if (grad) {
        for (size_t k = 0; k < nc; k++) {
            const int g = cfg->group[k];
            M2D_CHECK(g < cfg->n_groups, M2D_ERR_BAD_ARG, "m2d_adjoint_energy_2d: group %d out of range", g);
            M2D_CHECK(g < 0 || f->h_mask[k] == M2D_FIXED, M2D_ERR_BAD_ARG,
                      "m2d_adjoint_energy_2d: grouped node %zu is not a FIXED electrode node", k);
        }
    }
    m2d_species2d sp = *s;
    if (!mcc) sp.lifetime = INFINITY;
    const uint32_t fl_fwd = M2D_STEP_PUSH | M2D_STEP_BOUNDARY | M2D_STEP_DEPOSIT;
    // ---- forward: rho^0, then solve / reset / step, storing u^n and checkpoints
    const auto t0 = clock::now();
    DBuf U, ckpt;
    const int nck = (N + K - 1) / K;
    if (grad) {
        M2D_TRY(U.alloc(sizeof(double) * nc * (size_t)N));
        M2D_TRY(ckpt.alloc(sizeof(float) * 5 * np * (size_t)nck));
    }
    float *const parr[5] = {p->x, p->z, p->vx, p->vy, p->vz};
    M2D_TRY(m2d_fields2d_reset(f));
    M2D_TRY(m2d_accumulate_2d(ctx, &f->p.grid, p, s->charge, f->rho));
    m2d_solve_info si;
    for (int n = 0; n < N; n++) {
        if (grad && n % K == 0)
            for (int c = 0; c < 5; c++)
                M2D_HIP(hipMemcpyAsync(ckpt.as<float>() + ((size_t)(n / K) * 5 + c) * np, parr[c],
                                       sizeof(float) * np, hipMemcpyDeviceToDevice, st));
        M2D_TRY(m2d_fields2d_solve(f, M2D_SOLVE_STATIC, 1e-10, &si));
        if (grad)
            M2D_HIP(hipMemcpyAsync(U.as<double>() + (size_t)n * nc, f->u, sizeof(double) * nc,
                                   hipMemcpyDeviceToDevice, st));
        M2D_TRY(m2d_fields2d_reset(f));
        sp.niter = s->niter + (uint64_t)n;
        M2D_TRY(m2d_advance_2d(ctx, f, &sp, p, mcc, f->rho, fl_fwd, nullptr));
    }
    double mean = 0.0;
    int64_t live = 0;
    M2D_TRY(m2d_mean_energy_2d(ctx, p, s->mass, &mean, &live));
    *J = mean;
    const auto t1 = clock::now();
    if (info) {
        info->n_live = live;
        info->forward_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        info->backward_ms = 0.0;
    }
    if (!grad) return M2D_OK;
    M2D_CHECK(live > 0, M2D_ERR_BAD_ARG, "m2d_adjoint_energy_2d: no particle survived");

    // ---- reverse sweep
    DBuf lam, lu, lrho, seg, rep, gpart, dgrad, dgroup;
    M2D_TRY(lam.alloc(sizeof(double) * 5 * np));
    M2D_TRY(lu.alloc(sizeof(double) * nc));
    M2D_TRY(lrho.alloc(sizeof(double) * nc));
    // per replayed step: x, z, and with MCC the velocities vx, vy, vz
    const int SC = mcc ? 5 : 2;
    M2D_TRY(seg.alloc(sizeof(float) * SC * np * (size_t)(K + 1)));
    M2D_TRY(rep.alloc(sizeof(float) * 5 * np));
    const unsigned gb = nblk(nc);
    M2D_TRY(gpart.alloc(sizeof(double) * (size_t)gb * cfg->n_groups));
    M2D_TRY(dgrad.alloc(sizeof(double) * cfg->n_groups));
    M2D_TRY(dgroup.alloc(sizeof(int) * nc));
    M2D_HIP(hipMemcpy(dgroup.p, cfg->group, sizeof(int) * nc, hipMemcpyHostToDevice));
    M2D_TRY(fill_bytes(ctx, dgrad.p, 0, sizeof(double) * cfg->n_groups));
    M2D_TRY(fill_bytes(ctx, lrho.p, 0, sizeof(double) * nc));
    double *lx = lam.as<double>(), *lz = lx + np, *lvx = lz + np, *lvy = lvx + np, *lvz = lvy + np;
    k_adj_init2d<<<nblk(np), 256, 0, st>>>(p->x, p->vx, p->vy, p->vz, p->n,
                                           s->mass / (kQE * (double)live), lx, lz, lvx, lvy, lvz);
    M2D_HIP(hipGetLastError());
    // the adjoint solver: same mask and operator, zero Dirichlet values
    m2d_fields2d *af = nullptr;
    {
        std::vector<double> zero(nc, 0.0);
        M2D_TRY(m2d_fields2d_create(ctx, &f->p, f->h_mask.data(), zero.data(), &af));
    }
    struct Closer { m2d_fields2d *f; ~Closer() { m2d_fields2d_destroy(f); } } closer{af};
    AdjArgs a{};
    a.n = p->n;
    a.lx = lx; a.lz = lz; a.lvx = lvx; a.lvy = lvy; a.lvz = lvz;
    a.lrho1 = lrho.as<double>();
    a.lu = lu.as<double>();
    a.nx = f->nx; a.nz = f->nz;
    a.idx = f->idx; a.idz = f->idz;
    a.xmin = f->p.grid.xmin; a.zmin = f->p.grid.zmin;
    a.q = s->charge; a.dt = s->dt;
    a.c = s->charge * s->dt / (2.0 * s->mass);
    {
        // the rotation of advance_boris (particles.cpp:959-979) as a matrix, applied to unit vectors
        const double tmp = s->charge * s->dt / (2.0 * s->mass);
        const double tx = f->p.Br * tmp, ty = f->p.Bt * tmp, tz = f->p.Bz * tmp;
        const double sc = 2.0 / (1 + tx * tx + ty * ty + tz * tz);
        const double sx = tx * sc, sy = ty * sc, sz = tz * sc;
        for (int col = 0; col < 3; col++) {
            double v[3] = {0, 0, 0};  // (vx, vy, vz)
            v[col] = 1.0;
            const double vx = v[0], vy = v[1], vz = v[2];
            const double vpr = vx - vy * tz + vz * ty, vpz = vz - vx * ty + vy * tx, vpt = vy - vz * tx + vx * tz;
            a.M[0 * 3 + col] = vx - vpt * sz + vpz * sy;
            a.M[1 * 3 + col] = vy - vpz * sx + vpr * sz;
            a.M[2 * 3 + col] = vz - vpr * sy + vpt * sx;
        }
    }
    a.counters = ctx->counters;
    if (mcc) {
        a.mcc = mcc->d;
        a.thr = trial_threshold(1.0 - exp(-s->dt / s->lifetime));  // the pass's own threshold
        a.seed = ctx->seed;
        a.stream = s->stream;
    }
    const double rscale = -(f->p.grid.dx * f->p.grid.dx) / kEps0 / f->p.dV;  // b_F = rscale * rho
    float *const rarr[5] = {rep.as<float>(), rep.as<float>() + np, rep.as<float>() + 2 * np,
                            rep.as<float>() + 3 * np, rep.as<float>() + 4 * np};
    const m2d_particles2d rp{rarr[0], rarr[1], rarr[2], rarr[3], rarr[4], p->n};
    float *segx = seg.as<float>();
    for (int s0 = ((N - 1) / K) * K; s0 >= 0; s0 -= K) {
        const int s1 = std::min(N, s0 + K);
        // replay the segment's pushes from its checkpoint with the stored u^n
        for (int c = 0; c < 5; c++)
            M2D_HIP(hipMemcpyAsync(rarr[c], ckpt.as<float>() + ((size_t)(s0 / K) * 5 + c) * np, sizeof(float) * np,
                                   hipMemcpyDeviceToDevice, st));
        // snapshot k of the segment: the components of the store before step s0 + k
        auto snap = [&](int k) -> int {
            float *o = segx + (size_t)k * SC * np;
            for (int c = 0; c < SC; c++)
                M2D_HIP(hipMemcpyAsync(o + (size_t)c * np, rarr[c], sizeof(float) * np,
                                       hipMemcpyDeviceToDevice, st));
            return M2D_OK;
        };
        M2D_TRY(snap(0));
        for (int n = s0; n < s1; n++) {
            M2D_HIP(hipMemcpyAsync(f->u, U.as<double>() + (size_t)n * nc, sizeof(double) * nc,
                                   hipMemcpyDeviceToDevice, st));
            M2D_TRY(fields_build_tables(f));
            sp.niter = s->niter + (uint64_t)n;
            M2D_TRY(m2d_advance_2d(ctx, f, &sp, &rp, mcc, nullptr, M2D_STEP_PUSH | M2D_STEP_BOUNDARY, nullptr));
            M2D_TRY(snap(n - s0 + 1));
        }
        // reverse steps of the segment
        for (int n = s1 - 1; n >= s0; n--) {
            const float *xa = segx + (size_t)(n - s0) * SC * np, *xb = segx + (size_t)(n - s0 + 1) * SC * np;
            a.x0 = xa; a.z0 = xa + np;
            a.x1 = xb; a.z1 = xb + np;
            a.u = U.as<double>() + (size_t)n * nc;
            M2D_TRY(fill_bytes(ctx, lu.p, 0, sizeof(double) * nc));
            if (mcc) {
                a.v0x = xa + 2 * np; a.v0y = xa + 3 * np; a.v0z = xa + 4 * np;
                a.step = s->niter + (uint64_t)n;
                k_adj_scatter2d<<<nblk(np), 256, 0, st>>>(a);
                M2D_HIP(hipGetLastError());
            }
            k_adj_step2d<<<nblk(np), 256, 0, st>>>(a);
            M2D_HIP(hipGetLastError());
            // A_FF y = lu_F
            k_adj_rhs<<<gb, 256, 0, st>>>(f->mask, lu.as<double>(), af->rho, nc);
            M2D_HIP(hipGetLastError());
            M2D_TRY(m2d_fields2d_solve(af, M2D_SOLVE_RAW, adj_tol, &si));
            k_adj_groups<<<gb, 256, 0, st>>>(f->mask, dgroup.as<int>(), lu.as<double>(), af->u, f->nx, f->nz,
                                             cfg->n_groups, gpart.as<double>());
            k_adj_groups_sum<<<cfg->n_groups, 256, 0, st>>>(gpart.as<double>(), (int)gb, cfg->n_groups,
                                                             dgrad.as<double>());
            k_adj_lrho<<<gb, 256, 0, st>>>(f->mask, af->u, rscale, lrho.as<double>(), nc);
            M2D_HIP(hipGetLastError());
        }
    }
    // leave the fields as the forward run left them
    M2D_HIP(hipMemcpyAsync(f->u, U.as<double>() + (size_t)(N - 1) * nc, sizeof(double) * nc,
                           hipMemcpyDeviceToDevice, st));
    M2D_TRY(fields_build_tables(f));
    M2D_HIP(hipMemcpyAsync(grad, dgrad.p, sizeof(double) * cfg->n_groups, hipMemcpyDeviceToHost, st));
    M2D_HIP(hipStreamSynchronize(st));
    if (info) info->backward_ms = std::chrono::duration<double, std::milli>(clock::now() - t1).count();
    return M2D_OK;
}

}  // extern "C"
