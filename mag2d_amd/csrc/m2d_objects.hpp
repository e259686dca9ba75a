// m2d_objects.hpp -- device-side objects shared by the kernels of
// libmag2d_amd: the Fields object (gather tables, cell mask, multigrid
// hierarchy) and the MCC tables.
#pragma once
#include "m2d_internal.hpp"

namespace m2d {

// ------------------------------------------------------------------ MCC
constexpr int kMaxTargets = 8;
constexpr int kMaxInter = 32;

struct McInter {
    int type;
    int tab_off, tab_len;
    int pad;
    double DE;      // J
    double rate;    // m^3/s, LANGEVIN already scaled by cutoff^2 (particles.hpp:81-82)
    double cutoff;
    double mu;      // reduced mass of primary and target
    // COULOMB (Interaction::coulomb_sigma, particles.cpp:19-26): the primary's
    // Debye length and q1 q2 / (4 pi eps0)
    double lambda_D, qq;
};

struct McTarget {
    double rate;    // rates_by_species (particles.cpp:163)
    double density;
    double mass;
    double vth;     // sqrt(k_B T / m) = v_max/sqrt(2) (particles.hpp:180,190-192)
    int first, n;   // interactions [first, first+n)
    const float *px, *pz;                    // partner ensemble (px == nullptr: continuum)
    float *pvx, *pvy, *pvz;                  // its velocities (COULOMB writes the partner's)
    int64_t pn;
};

struct McDev {
    int n_targets;
    int partner_update;  // some interaction (COULOMB) changes the partner's velocity
    double lifetime;
    double mass;    // primary
    McTarget t[kMaxTargets];
    McInter in[kMaxInter];
    const double *tabE;
    const double *tabS;
};

// ------------------------------------------------------------------ multigrid
struct MgLevel {
    int nx, nz;
    size_t n;
    float *st;      // 9 coefficients per node, offsets (di,dj) in row-major (di+1)*3+(dj+1)
    float *dinv;    // inverse diagonal, 0 on Dirichlet nodes
    uint8_t *dir;   // 1 = correction pinned to 0
    float *x, *x2, *r, *b;
};

struct Mg {
    std::vector<MgLevel> lv;
    float *coarse_inv = nullptr;  // dense inverse of the coarsest operator (n x n)
    float *coarse_invT = nullptr; // the same, transposed
    int coarse_n = 0;
    // FP32 PCG vectors on the finest level
    float *pr = nullptr, *pz = nullptr, *pp = nullptr, *pp2 = nullptr, *pq = nullptr, *pe = nullptr;
    // FP64 vectors
    double *r64 = nullptr;
    double *dots = nullptr;       // device scalars
    double *h_dots = nullptr;     // pinned mirror
    double *partial = nullptr;    // reduction partials
    int npartial = 0;
    unsigned int *ticket = nullptr;  // block counter of the fused reductions (0 between launches)
    int lc = 0;                   // first level of the one-workgroup coarse V-cycle
    SolveEst est;  // the last solve's passes and iterations: what the next queues
};

}  // namespace m2d

struct m2d_fields2d {
    m2d_ctx *ctx = nullptr;
    m2d_fields2d_params p{};
    double idx = 0, idz = 0;
    int nx = 0, nz = 0;
    size_t n = 0;
    uint8_t *mask = nullptr;      // device copy of t_grid::mask
    uint8_t *cellfree = nullptr;  // nx*nz: 1 if any corner of cell (i,j) is FREE
    uint8_t *winfree = nullptr;   // per sort tile: 1 if every cell of its deposit window is free
    double *voltage = nullptr, *u = nullptr, *urf = nullptr, *rho = nullptr;
    float2 *gx = nullptr;         // (nx+1)*nz edge x-gradients {u, uRF}, padded rows 0 and nx
    float2 *gz = nullptr;         // nx*(nz+1) edge z-gradients {u, uRF}, padded cols 0 and nz
    // gx, gz with one step's RF phase applied, fmaf(phase, uRF part, u part)
    // per edge: written by the flush after a pass for the next step's phase,
    // read by the specialised tiled pass (one float per edge sample instead
    // of two).  Valid while ph_valid, for the FP32 phase ph_phase;
    // fields_build_tables invalidates them.
    float *phx = nullptr, *phz = nullptr;
    float ph_phase = 0.f;
    bool ph_valid = false;
    // gridded B (Fields::Br, Bz with magnetic_field_const = 0): {Br, Bz} per
    // node of its own grid, null for the constant B of the params
    float2 *bt = nullptr;
    m2d_grid2d bgrid{};
    double *smooth_tmp = nullptr;  // nx*nz copy of the field u_smooth filters
    m2d::Mg mg;
    std::vector<uint8_t> h_mask;
};

struct m2d_mcc {
    m2d_ctx *ctx = nullptr;
    m2d::McDev h{};               // host copy
    m2d::McDev *d = nullptr;      // device copy
    double *tab = nullptr;        // device cross-section tables (E then sigma)
    std::vector<double> rates;
};

namespace m2d {
int fields_build_tables(m2d_fields2d *f);
int mcc_sync(m2d_mcc *m);
// some target draws its partners from the store whose first position array
// is x (the colliding species itself)
inline bool mcc_partners_alias(const m2d_mcc *m, const void *x)
{
    for (int t = 0; t < m->h.n_targets; t++)
        if ((const void *)m->h.t[t].px == x) return true;
    return false;
}
}  // namespace m2d
