// m2d_particles2d.hip -- the 2D PIC-MCC particle pass and its helpers.
//
// k_advance2d fuses Species<CARTESIAN>::advance (particles.hpp:342-349):
//   advance_boris (particles.cpp:924-995): E gather (Fields::E, fields.hpp:124-150
//   on Field2D::grad, Field2D.hpp:80-168), Boris kick-rotate-kick, drift, MCC
//   trial and BaseSpecies::scatter (particles.cpp:208-365);
//   advance_boundary (particles.hpp:370-411): box test, is_free, CIC deposit.
// One read and one write of the 20-byte SoA state per live particle.
#include "m2d_objects.hpp"
#include "m2d_scatter.hpp"
#include "m2d_gather2d.hpp"
#include <cmath>

namespace m2d {

// ------------------------------------------------------------------ args
struct PushArgs {
    float *x, *z, *vx, *vy, *vz;
    int64_t n;
    float *xl, *zl;                // position remainders (long-time positions), or null
    const float2 *gx, *gz;
    const float *phx, *phz;        // gx, gz with this step's phase applied (PHT pass)
    const uint8_t *cellfree;
    const uint8_t *winfree;        // per sort tile: whole deposit window free
    int nx, nz;
    double idx, idz, xmin, zmin;
    float idxh, idxl, idzh, idzl;  // idx, idz as FP32 pairs (the origin-0 gather)
    // box [x_min,x_max] x [z_min,z_max] as the FP32 bounds that give the same
    // answer for FP32 positions; x_max, z_max in FP64 for the periodic wrap
    float fx0, fx1, fz0, fz1;
    double xper, zper;
    float qmdt2;                   // q dt / (2 m)
    float tx, ty, tz, sx, sy, sz;  // Boris rotation vectors
    float dt;
    float dtl;                     // dt - (double)dt as FP32: the drift of long-time positions
    float phase;                   // E = -(grad u + phase grad uRF)
    float ecx, ecz;                // constant field (EMPTY geometry, fields.hpp:126-131)
    int econst;
    int boundary, periodic, checkfree;
    uint64_t trial_thr;            // trial word w fires iff w < trial_thr (uni() < 1 - exp(-dt/lifetime))
    int rotate;                    // B != 0
    int push;                      // advance_boris; 0: boundary and deposit only
    uint64_t seed;
    uint32_t stream;
    uint64_t step;
    const McDev *mcc;
    uint32_t *cq;                  // collision queue: slots whose MCC trial fired
    float4 *cqv;                   // the queued slot's velocity after the push (vx, vy, vz, 0)
    uint32_t cq_cap;               // capacity of both in entries
    int cq_defer;                  // scatter results go back to cqv; k_scatter_apply stores them
    unsigned long long *cq_len;    // its length (one of two counters, alternating per launch)
    unsigned long long *acc;       // nx*nz fixed-point deposit grid
    unsigned long long *counters;
    // tiled deposit (store sorted by m2d_sort_2d)
    const int4 *sched;             // chunk -> (tile, first slot, end slot)
    const unsigned int *n_chunks;
    int ntz;                       // sort tiles along z
    // gridded B (Fields::B with magnetic_field_const = 0): {Br, Bz} per node
    // of its own grid, null for the constant rotation above
    const float2 *bt;
    int bnx, bnz;
    double bidx, bidz, bxmin, bzmin;
    double rscale;                 // q dt / (2 m) times the rotation's scale
};

// Fields::B with magnetic_field_const = 0 (fields.hpp:171-176): Br, Bz by
// Field2D::interpolate (Field2D.hpp:64-78) of the B grid, Bt = 0.  Index
// arithmetic in FP64, weights FP32.  Returns true for a point outside the B
// grid (a NaN included): it is clamped onto the grid, where the reference
// throws.
__device__ __forceinline__ bool gather_B(const PushArgs &a, double xd, double zd, float &br, float &bz)
{
    double X = (xd - a.bxmin) * a.bidx, Z = (zd - a.bzmin) * a.bidz;
    const bool bad = !(X >= 0.0 && X <= a.bnx - 1.0 && Z >= 0.0 && Z <= a.bnz - 1.0);
    if (bad) {  // fmax of a NaN and 0 is 0
        X = fmin(fmax(X, 0.0), (double)(a.bnx - 1));
        Z = fmin(fmax(Z, 0.0), (double)(a.bnz - 1));
    }
    const int i = min((int)X, a.bnx - 2), j = min((int)Z, a.bnz - 2);
    const float u = (float)(X - i), v = (float)(Z - j);
    const float2 *r0 = a.bt + ((size_t)i * a.bnz + j), *r1 = r0 + a.bnz;
    const float2 p00 = r0[0], p01 = r0[1], p10 = r1[0], p11 = r1[1];
    const float w00 = (1.f - u) * (1.f - v), w10 = u * (1.f - v), w01 = (1.f - u) * v, w11 = u * v;
    br = w00 * p00.x + w10 * p10.x + w01 * p01.x + w11 * p11.x;
    bz = w00 * p00.y + w10 * p10.y + w01 * p01.y + w11 * p11.y;
    return bad;
}

// Boris rotation vectors t = B q dt/(2m), s = 2t/(1 + t^2) (particles.cpp:957-975)
// of the local field, with (Bx, By, Bz) = (Br, Bt, Bz) and Bt = 0
__device__ __forceinline__ bool local_rotation(const PushArgs &a, float x, float z, float &tx, float &ty, float &tz,
                                               float &sx, float &sy, float &sz)
{
    float br, bz;
    const bool bad = gather_B(a, (double)x, (double)z, br, bz);
    const double t1 = (double)br * a.rscale, t3 = (double)bz * a.rscale;
    const double sc = 2.0 / (1.0 + t1 * t1 + t3 * t3);
    tx = (float)t1; ty = 0.f; tz = (float)t3;
    sx = (float)(t1 * sc); sy = 0.f; sz = (float)(t3 * sc);
    return bad;
}

// ------------------------------------------------------------------ deposit
// CIC weights (Field2D.hpp:45-62) as 2^-31 fixed point.  The cell offsets
// u, v and their complements 1 - u, 1 - v are formed in FP64 from the
// products x*idx (exact subtraction of the cell index) and rounded to FP32
// each, so every factor carries a relative error of at most 2^-24 (a
// complement formed in FP32 from a rounded u would lose its relative
// accuracy as u nears 1); the four bilinear products are FP32 (relative
// error ~2e-7 in all), and the rounded weights are corrected on the nearest
// node's (at least 1/4) so each particle deposits exactly 2^31: charge is
// conserved bit for bit and node sums do not depend on the order of the
// particles.
// NONNEG: the caller guarantees X, Z >= 0 (a NaN included: it fails the
// upper bounds).
// The four weights from the offsets u, v and their complements ou, ov.
__device__ __forceinline__ void cic_from_offsets(float u, float ou, float v, float ov, uint32_t w[4])
{
    // scaling by 2^31 is exact, so ou * (ov 2^31) has the bits of (ou ov) 2^31
    const float S = 2147483648.0f;  // 2^31
    const float a = ov * S, b = v * S;
    w[0] = __float2uint_rn(ou * a);
    w[1] = __float2uint_rn(u * a);
    w[2] = __float2uint_rn(ou * b);
    w[3] = __float2uint_rn(u * b);
    // each weight is at most 2^31, so the sum stays below 2^32
    const uint32_t d = 0x80000000u - (w[0] + w[1] + w[2] + w[3]);
    // the correction goes to the weight of the node nearest the point, which
    // is at least 1/4 (the largest, up to a rounding at u or v = 1/2)
    const bool bu = u >= 0.5f, bv = v >= 0.5f;
    w[0] += !bu && !bv ? d : 0u;  // modulo 2^32: a few units of either sign
    w[1] += bu && !bv ? d : 0u;
    w[2] += !bu && bv ? d : 0u;
    w[3] += bu && bv ? d : 0u;
}

template <bool NONNEG = false>
__device__ __forceinline__ bool cic_weights(double X, double Z, int nx, int nz, int &i, int &j,
                                            uint32_t w[4])
{
    // range in FP64 before the conversion: cells [0, n-2]
    if (!((NONNEG || X >= 0.0) && X < nx - 1.0 && (NONNEG || Z >= 0.0) && Z < nz - 1.0)) return false;
    i = (int)X;
    j = (int)Z;
    const double du = X - i, dv = Z - j;
    cic_from_offsets((float)du, (float)(1.0 - du), (float)dv, (float)(1.0 - dv), w);
    return true;
}


__device__ __forceinline__ void deposit_global(unsigned long long *acc, int nz, int i, int j,
                                               const long long w[4])
{
    size_t k = (size_t)i * nz + j;
    if (w[0]) atomicAdd(acc + k, (unsigned long long)w[0]);
    if (w[1]) atomicAdd(acc + k + nz, (unsigned long long)w[1]);
    if (w[2]) atomicAdd(acc + k + 1, (unsigned long long)w[2]);
    if (w[3]) atomicAdd(acc + k + nz + 1, (unsigned long long)w[3]);
}

// ------------------------------------------------------------------ the pass
// A block of kPass threads owns a chunk of consecutive slots, taken in rounds
// of 4 kPass.  Thread t holds slots 4t..4t+3 of a round in registers: each
// component moves as one aligned float4 per thread (a wavefront moves 1 KB
// contiguous per component), and consecutive particles of a sorted store,
// which mostly share a cell, meet in one thread.
// Deposits: the tiled pass adds every particle's four weights into its
// block's LDS window (in a sorted store some lane of a wavefront changes cell
// at almost every particle, so summing runs first would not save LDS
// instructions, only add VALU work).  The untiled pass sums them twice before
// they reach memory: in registers over a thread's run of equal cells, then
// across the wavefront by a segmented scan over the lanes' last runs (lanes
// hold consecutive slots, so a wavefront of a sorted store spans a few
// cells); only the last lane of each segment issues the four 64-bit atomics.
// Two-wavefront blocks with 5 KB of LDS (the deposit window): a launch of
// a few thousand chunks is mostly resident at once.
constexpr int kPass = 256;
constexpr int kRound = kPass * 4;
static_assert(kRound == kRound2, "the tile schedule orders chunks by rounds of the pass");
constexpr int kGather = 2;                  // particles whose gathers are in flight together
constexpr int kChunk = 2048;                // slots per block of the untiled pass
constexpr int kPPT = kChunk / kBlock;       // slots per thread of the helper kernels

// component c of the state (x, z, vx, vy, vz, and the position remainders
// xl, zl of long-time positions); c is a constant after unrolling, so the
// pointer stays in scalar registers (an array of the pointers would live in
// scratch memory)
__device__ __forceinline__ float *comp(const PushArgs &a, int c)
{
    return c == 0 ? a.x : c == 1 ? a.z : c == 2 ? a.vx : c == 3 ? a.vy : c == 4 ? a.vz : c == 5 ? a.xl : a.zl;
}

// running sum of one thread's deposits while consecutive particles share a cell
struct DepRun {
    int ci, cj;
    long long w[4];
};

// one particle's weights into the block's LDS window (tile + halo), or
// into the global grid when its cell lies outside the window
__device__ __forceinline__ void window_deposit(int ci, int cj, const uint32_t w[4], unsigned long long *win,
                                               int wi0, int wj0, unsigned long long *acc, int nz)
{
    const int a = ci - wi0, b = cj - wj0;
    if (a >= 0 && a + 1 < kWin && b >= 0 && b + 1 < kWin) {
        unsigned long long *s = win + a * kWin + b;
        atomicAdd(s, (unsigned long long)w[0]);
        atomicAdd(s + kWin, (unsigned long long)w[1]);
        atomicAdd(s + 1, (unsigned long long)w[2]);
        atomicAdd(s + kWin + 1, (unsigned long long)w[3]);
    } else {
        const long long ww[4] = {w[0], w[1], w[2], w[3]};
        deposit_global(acc, nz, ci, cj, ww);
    }
}

// segmented inclusive scan of the lanes' runs keyed by cell; segment tails
// deposit.  All 64 lanes must call it together.
__device__ __forceinline__ void wave_deposit(DepRun &d, unsigned long long *acc, int nz)
{
    const int lane = __lane_id();
    const bool valid = d.ci >= 0;
    const int key = valid ? d.ci * nz + d.cj : -1 - lane;
    const int prev = __shfl_up(key, 1);
    const unsigned long long heads = __ballot(lane == 0 || prev != key);
    const int seg = 63 - __clzll(heads & ((2ull << lane) - 1ull));
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        long long u0 = __shfl_up(d.w[0], off), u1 = __shfl_up(d.w[1], off);
        long long u2 = __shfl_up(d.w[2], off), u3 = __shfl_up(d.w[3], off);
        if (lane - off >= seg) { d.w[0] += u0; d.w[1] += u1; d.w[2] += u2; d.w[3] += u3; }
    }
    const bool tail = lane == 63 || ((heads >> (lane + 1)) & 1ull);
    if (valid && tail) deposit_global(acc, nz, d.ci, d.cj, d.w);
    d.ci = -1;
}

// FAST: the common configuration fixed at compile time (B = 0, no periodic
// wrap, no constant field, boundary and free-cell tests on, origin at 0),
// which frees the scalar registers and branches the generic flags cost.
// BGRID (generic pass only): B from the gridded field at each particle.
// PHT (FAST only): the gather reads the phase-applied tables phx, phz.
// Wavefronts per SIMD each form is built for: a thread holds its four
// particles and two gathers in registers.  The specialised tiled pass's FP32
// index arithmetic (gather, cell and CIC offsets) fits six wavefronts (80
// VGPRs) when the gather reads the one-float phase-applied tables, five (96
// VGPRs) with the two-component tables; the untiled forms carry the deposit
// runs and the FP64 generic forms their wider arithmetic, and keep four (128
// VGPRs), three for the untiled deposit with a gridded B; with long-time
// positions (two more state components) the forms keep four, three for the
// untiled generic deposit.  Each is the most that builds without spilling to
// scratch memory, which costs far more than the occupancy gains.
template <bool DEPOSIT, bool TILED, bool FAST, bool BGRID, bool PHT, bool HIPOS>
constexpr int pass_waves()
{
    if (HIPOS) return !TILED && DEPOSIT && !FAST ? 3 : 4;  // two more state components
    return TILED && FAST ? (PHT ? 6 : 5) : (BGRID && DEPOSIT && !TILED ? 3 : 4);
}

// HIPOS: long-time positions, the FP32 pairs (x, xl), (z, zl).
template <bool DEPOSIT, bool MCC, bool TILED, bool FAST, bool BGRID = false, bool PHT = false, bool HIPOS = false>
__global__ void __launch_bounds__(kPass)
__attribute__((amdgpu_waves_per_eu(pass_waves<DEPOSIT, TILED, FAST, BGRID, PHT, HIPOS>())))
k_advance2d(PushArgs a)
{
    __shared__ unsigned long long win[TILED ? kWin * kWin : 1];
    const int tid = threadIdx.x;
    // slots [lo, hi) of this block; rounds start at lo rounded down to 64, so
    // a thread's 4 slots are one aligned float4 of each component and share
    // one trial Philox block
    int64_t lo, hi, co0;
    int wi0 = 0, wj0 = 0;
    bool wfree = false;
    if constexpr (TILED) {
        // chunks in runs per XCD: neighbouring tiles share an L2
        const unsigned nch = *a.n_chunks;
        const unsigned ch = xcd_run_block(blockIdx.x);
        if (ch >= nch) return;  // uniform: before any barrier
        const int4 s = a.sched[ch];
        lo = s.y;
        hi = s.z;
        co0 = lo & ~(int64_t)63;
        wi0 = (s.x / a.ntz) * kTile - kHalo;
        wj0 = (s.x % a.ntz) * kTile - kHalo;
        wfree = a.winfree[s.x] != 0;
        for (int k = tid; k < kWin * kWin; k += kPass) win[k] = 0ull;
        __syncthreads();
    } else {
        lo = (int64_t)blockIdx.x * kChunk;
        hi = min(lo + (int64_t)kChunk, a.n);
        co0 = lo;
    }
    int n_removed = 0, n_coll = 0;
    bool nan_any = false, gbad = false;
    DepRun run;
    run.ci = -1;
#pragma unroll 1
    for (int64_t co = co0; co < hi; co += kRound) {  // uniform over the block
        // ---- this thread's slots pb..pb+3; slots outside [lo, hi) belong to
        // other blocks: they are neither computed nor written
        const int64_t pb = co + 4 * tid;
        const bool full = pb >= lo && pb + 4 <= hi;
        constexpr int NC = HIPOS ? 7 : 5;  // state components held
        float v[7][4];
        if (full) {
#pragma unroll
            for (int c = 0; c < NC; c++) {
                const float4 t = *reinterpret_cast<const float4 *>(comp(a, c) + pb);
                v[c][0] = t.x; v[c][1] = t.y; v[c][2] = t.z; v[c][3] = t.w;
            }
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (pb + k >= lo && pb + k < hi) {
#pragma unroll
                    for (int c = 0; c < NC; c++) v[c][k] = comp(a, c)[pb + k];
                } else {
                    v[0][k] = dead_value();
#pragma unroll
                    for (int c = 1; c < NC; c++) v[c][k] = 0.f;
                }
            }
        }
        uint32_t tw[4] = {0u, 0u, 0u, 0u};
        if (MCC) {
            // trial uniforms of slots pb..pb+3: one Philox block
            tw[0] = (uint32_t)(pb >> 2); tw[1] = (uint32_t)a.step; tw[2] = 0u;
            tw[3] = (kPurposeTrial << 24) | ((uint32_t)(a.step >> 32) & 0xFFFFFFu);
            philox4x32_10(tw[0], tw[1], tw[2], tw[3], (uint32_t)a.seed, (uint32_t)(a.seed >> 32) ^ a.stream);
        }
        // ---- particles in groups of kGather: the group's gathers are issued
        // before any is used, so kGather table fetches are in flight per
        // thread.  The body works on entries 0..kGather-1 and the arrays
        // rotate by kGather after each group (register renaming once the
        // loop is unrolled).
#pragma unroll
        for (int q0 = 0; q0 < 4; q0 += kGather) {
            float gex[kGather], gez[kGather];
#pragma unroll
            for (int h = 0; h < kGather; h++) gex[h] = gez[h] = 0.f;
            if (FAST || a.push) {
#pragma unroll
                for (int h = 0; h < kGather; h++) {
                    // a dead slot gathers at a clamped point and is discarded
                    bool b;
                    if constexpr (FAST)
                        b = gather_E_f<PHT>(a, v[0][h], HIPOS ? v[5][h] : 0.f, v[1][h], HIPOS ? v[6][h] : 0.f,
                                            a.phase, gex[h], gez[h]);
                    else
                        b = gather_E_clamped<false, PHT>(a, HIPOS ? pos64(v[0][h], v[5][h]) : (double)v[0][h],
                                                         HIPOS ? pos64(v[1][h], v[6][h]) : (double)v[1][h], a.phase,
                                                         gex[h], gez[h]);
                    gbad |= b && !is_dead(v[0][h]);
                }
            }
#pragma unroll
            for (int q = 0; q < kGather; q++) {
                const int h = q;
                float x = v[0][q];
                if (is_dead(x)) continue;
                float z = v[1][q], vx = v[2][q], vy = v[3][q], vz = v[4][q];
                float xl = HIPOS ? v[5][q] : 0.f, zl = HIPOS ? v[6][q] : 0.f;

                if (FAST || a.push) {
                    // ---- advance_boris (particles.cpp:945-988)
                    const float ex = gex[h], ez = gez[h];
                    vx = fmaf(ex, a.qmdt2, vx);
                    vz = fmaf(ez, a.qmdt2, vz);
                    if (!FAST && a.rotate) {  // with B = 0 the rotation is the identity, bit for bit
                        float tx = a.tx, ty = a.ty, tz = a.tz, sx = a.sx, sy = a.sy, sz = a.sz;
                        // B at the position of the E gather, before the drift
                        if (BGRID) gbad |= local_rotation(a, x, z, tx, ty, tz, sx, sy, sz);
                        float vpr = vx - vy * tz + vz * ty;
                        float vpz = vz - vx * ty + vy * tx;
                        float vpt = vy - vz * tx + vx * tz;
                        vx = vx - vpt * sz + vpz * sy;
                        vy = vy - vpz * sx + vpr * sz;
                        vz = vz - vpr * sy + vpt * sx;
                    }
                    vx = fmaf(ex, a.qmdt2, vx);
                    vz = fmaf(ez, a.qmdt2, vz);
                    if constexpr (HIPOS) {
                        drift_pair(x, xl, vx, a.dt, a.dtl);
                        drift_pair(z, zl, vz, a.dt, a.dtl);
                    } else {
                        x = fmaf(vx, a.dt, x);
                        z = fmaf(vz, a.dt, z);
                    }

                    // ---- MCC trial (particles.cpp:990-993).  A hit only changes the
                    // velocity after the drift, so it cannot change the boundary test
                    // or the deposit below: the slot is queued and k_flush_scatter
                    // applies BaseSpecies::scatter afterwards.
                    if (MCC) {
                        if ((uint64_t)tw[q] < a.trial_thr) {
                            unsigned long long at = atomicAdd(a.cq_len, 1ull);
                            a.cq[at] = (uint32_t)(pb + q0 + q);
                            a.cqv[at] = make_float4(vx, vy, vz, 0.f);
                            n_coll++;
                        }
                    }
                }

                // ---- advance_boundary (particles.hpp:376-406)
                bool dead = false;
                if (FAST) {
                    // a NaN fails every comparison, so it lands outside; a
                    // long-time position is tested in FP64 against the box
                    if constexpr (HIPOS) {
                        const double xd = pos64(x, xl), zd = pos64(z, zl);
                        dead = !(xd >= 0.0 && xd <= a.xper && zd >= 0.0 && zd <= a.zper);
                    } else {
                        dead = !(x >= a.fx0 && x <= a.fx1 && z >= a.fz0 && z <= a.fz1);
                    }
                    nan_any |= dead && (x != x || z != z);
                } else if (a.boundary) {
                    const double xd = HIPOS ? pos64(x, xl) : (double)x, zd = HIPOS ? pos64(z, zl) : (double)z;
                    bool inside = HIPOS ? xd >= a.xmin && xd <= a.xper && zd >= a.zmin && zd <= a.zper
                                        : x >= a.fx0 && x <= a.fx1 && z >= a.fz0 && z <= a.fz1;
                    if (!inside) {
                        if (x != x || z != z) { dead = true; nan_any = true; }
                        else if (!a.periodic) dead = true;
                        else {
                            double xm = fmod(xd, a.xper), zm = fmod(zd, a.zper);
                            if (xm < 0) xm += a.xper;
                            if (zm < 0) zm += a.zper;
                            if constexpr (HIPOS) {
                                pos_split(xm, x, xl);
                                pos_split(zm, z, zl);
                            } else {
                                x = (float)xm; z = (float)zm;
                            }
                        }
                    }
                }
                // FAST: cell and offsets in FP32 pair arithmetic (the box
                // test above gave 0 <= x, z); otherwise FP64 cell coordinates
                int fci = 0, fcj = 0;
                float fu = 0.f, fou = 0.f, fv = 0.f, fov = 0.f;
                double X = 0.0, Z = 0.0;
                if constexpr (FAST) {
                    cell_f(x, xl, a.idxh, a.idxl, fci, fu, fou);
                    cell_f(z, zl, a.idzh, a.idzl, fcj, fv, fov);
                } else {
                    X = ((HIPOS ? pos64(x, xl) : (double)x) - a.xmin) * a.idx;
                    Z = ((HIPOS ? pos64(z, zl) : (double)z) - a.zmin) * a.idz;
                }
                if (!dead && (FAST || (a.boundary && a.checkfree))) {
                    bool free_ = false;
                    if (FAST || (X >= 0.0 && X < (double)a.nx && Z >= 0.0 && Z < (double)a.nz)) {
                        const int ci = FAST ? fci : (int)X, cj = FAST ? fcj : (int)Z;
                        // a cell of an all-free window needs no table lookup
                        if (wfree && (unsigned)(ci - wi0) < (unsigned)(kWin - 1) &&
                            (unsigned)(cj - wj0) < (unsigned)(kWin - 1))
                            free_ = true;
                        else if ((unsigned)ci <= (unsigned)(a.nx - 1) && (unsigned)cj <= (unsigned)(a.nz - 1))
                            free_ = a.cellfree[ci * a.nz + cj] != 0;
                    }
                    dead = !free_;
                }
                if (dead) {
                    // the slot keeps the step's velocity: a COULOMB partner
                    // update of its collision still reads it
                    v[0][q] = dead_value(); v[2][q] = vx; v[3][q] = vy; v[4][q] = vz;
                    n_removed++;
                    continue;
                }
                // ---- deposit (particles.hpp:408-409): tiled, straight into the
                // LDS window; untiled, summed per run of equal cells
                if (DEPOSIT) {
                    int ci, cj;
                    uint32_t w[4];
                    bool in_grid;
                    if constexpr (FAST) {
                        // cells [0, n-2]; the free-cell test above already
                        // removed the last row and column (not free)
                        ci = fci;
                        cj = fcj;
                        in_grid = (unsigned)ci <= (unsigned)(a.nx - 2) && (unsigned)cj <= (unsigned)(a.nz - 2);
                        if (in_grid) cic_from_offsets(fu, fou, fv, fov, w);
                    } else {
                        in_grid = cic_weights(X, Z, a.nx, a.nz, ci, cj, w);
                    }
                    if (in_grid) {
                        if constexpr (TILED) {
                            window_deposit(ci, cj, w, win, wi0, wj0, a.acc, a.nz);
                        } else if (ci == run.ci && cj == run.cj) {
                            run.w[0] += w[0]; run.w[1] += w[1]; run.w[2] += w[2]; run.w[3] += w[3];
                        } else {
                            if (run.ci >= 0) deposit_global(a.acc, a.nz, run.ci, run.cj, run.w);
                            run.ci = ci; run.cj = cj;
                            run.w[0] = w[0]; run.w[1] = w[1]; run.w[2] = w[2]; run.w[3] = w[3];
                        }
                    } else {
                        raise_flag(a.counters, M2D_FLAG_OUT_OF_RANGE);
                    }
                }
                v[0][q] = x; v[1][q] = z; v[2][q] = vx; v[3][q] = vy; v[4][q] = vz;
                if constexpr (HIPOS) { v[5][q] = xl; v[6][q] = zl; }
            }
#pragma unroll
            for (int c = 0; c < NC; c++) {
                float t[4];
#pragma unroll
                for (int q = 0; q < 4; q++) t[q] = v[c][(q + kGather) & 3];
#pragma unroll
                for (int q = 0; q < 4; q++) v[c][q] = t[q];
            }
            uint32_t u[4];
#pragma unroll
            for (int q = 0; q < 4; q++) u[q] = tw[(q + kGather) & 3];
#pragma unroll
            for (int q = 0; q < 4; q++) tw[q] = u[q];
        }
        if constexpr (DEPOSIT && !TILED) wave_deposit(run, a.acc, a.nz);
        // ---- store
        if (full) {
#pragma unroll
            for (int c = 0; c < NC; c++)
                *reinterpret_cast<float4 *>(comp(a, c) + pb) = make_float4(v[c][0], v[c][1], v[c][2], v[c][3]);
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (pb + k >= lo && pb + k < hi) {
#pragma unroll
                    for (int c = 0; c < NC; c++) comp(a, c)[pb + k] = v[c][k];
                }
            }
        }
    }
    if constexpr (TILED) {
        // the window's nonzero nodes go to the global grid as 64-bit atomics
        // (integer sums: the result does not depend on their order).  A
        // nonzero node belongs to a deposited cell, so it lies in the grid.
        __syncthreads();
        for (int k = tid; k < kWin * kWin; k += kPass) {
            const unsigned long long v = win[k];
            if (v) atomicAdd(a.acc + (size_t)(wi0 + k / kWin) * a.nz + (wj0 + k % kWin), v);
        }
    }
    // ---- counters: one atomic per wavefront
    int sum = n_removed;
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
    if (__lane_id() == 0 && sum) atomicAdd(&a.counters[CNT_REMOVED], (unsigned long long)sum);
    if (MCC) {
        int c = n_coll;
        for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
        if (__lane_id() == 0 && c) atomicAdd(&a.counters[CNT_COLL], (unsigned long long)c);
    }
    if (__ballot(nan_any) && __lane_id() == 0) raise_flag(a.counters, M2D_FLAG_NAN);
    if (__ballot(gbad) && __lane_id() == 0) raise_flag(a.counters, M2D_FLAG_OUT_OF_RANGE);
}

// BaseSpecies::scatter for the slots queued by k_advance2d, threads
// [start, start + stride, ...); each slot draws from its own Philox stream,
// so the queue order does not matter.  Thread 0 clears the other queue
// counter for the next launch.  The scatter is a chain of dependent reads
// and FP64 arithmetic that sets the flush launch's length, so the queue
// carries each slot's velocity (the particle's state is not read first), and
// the thread's first entry is read beside the queue length, not after it
// (the queue holds cq_cap entries: the read stays in bounds, and its value
// is used only below the length).  A slot that the boundary removed after
// its trial scatters too, as in the reference's loop (the scatter comes
// before advance_boundary there); nothing reads a removed slot's velocity.
__device__ __forceinline__ void scatter_queue(const PushArgs &a, unsigned long long start,
                                              unsigned long long stride, unsigned long long *next_len)
{
    const bool in_cap = start < a.cq_cap;
    const uint32_t p0 = in_cap ? a.cq[start] : 0u;
    const float4 w0 = in_cap ? a.cqv[start] : make_float4(0.f, 0.f, 0.f, 0.f);
    const unsigned long long len = *a.cq_len;
    if (start == 0) *next_len = 0ull;
    for (unsigned long long k = start; k < len; k += stride) {
        const uint32_t p = k == start ? p0 : a.cq[k];
        const float4 w = k == start ? w0 : a.cqv[k];
        float vx = w.x, vy = w.y, vz = w.z;
        Rng rg;
        rg.begin(a.seed, a.stream, p, a.step, kPurposeMcc);
        scatter_last(*a.mcc, rg, vx, vy, vz, a.counters);
        if (a.cq_defer) a.cqv[k] = make_float4(vx, vy, vz, 0.f);
        else { a.vx[p] = vx; a.vy[p] = vy; a.vz[p] = vz; }
    }
}

// The stores deferred by scatter_queue (cq_defer): when the partners come
// from the colliding species' own store, every scatter of the step reads
// its partner's velocity as the push left it, whatever the order in which
// the threads run, and only then do the new velocities land.
__global__ void k_scatter_apply(PushArgs a)
{
    const unsigned long long len = *a.cq_len;
    for (unsigned long long k = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; k < len;
         k += (unsigned long long)gridDim.x * blockDim.x) {
        const uint32_t p = a.cq[k];
        const float4 w = a.cqv[k];
        a.vx[p] = w.x; a.vy[p] = w.y; a.vz[p] = w.z;
    }
}

// BaseSpecies::scatter of the queued slots one after another in slot order,
// by one thread: an interaction that turns its partner too (COULOMB,
// particles.cpp:309-314) makes the order matter, and the reference scatters
// in slot order inside its particle loop.  Slots removed by the boundary
// after their trial scatter as well (the reference scatters before it
// removes), so their partners turn as there.  The thread also clears the
// other queue counter for the next launch.
__global__ void k_scatter_serial(PushArgs a, const uint32_t *sorted, unsigned long long len,
                                 unsigned long long *next_len)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    *next_len = 0ull;
    for (unsigned long long k = 0; k < len; k++) {
        const uint32_t p = sorted[k];
        float vx = a.vx[p], vy = a.vy[p], vz = a.vz[p];
        Rng rg;
        rg.begin(a.seed, a.stream, p, a.step, kPurposeMcc);
        scatter(*a.mcc, rg, vx, vy, vz, a.counters);
        a.vx[p] = vx; a.vy[p] = vy; a.vz[p] = vz;
    }
}

// The next step's phase-applied tables: phx[e] = fmaf(ph, gx[e].y, gx[e].x),
// the value the gather forms from gx at phase ph, and the same for phz
struct PhBuild {
    const float2 *gx, *gz;
    float *phx, *phz;
    size_t nxe, nze;  // entries of gx and gz
    float ph;
};
constexpr int kPhPerThread = 4;

// After the pass, in dispatch order: with MCC, kScatterBlocks blocks apply
// the collision queue (a long FP64 chain that sets the launch's length, so
// it starts first); ntab blocks build the next step's phase-applied tables;
// the rest add the deposit, rho += charge * acc * 2^-31, and clear acc.
constexpr unsigned kScatterBlocks = 64;
template <bool MCC>
__global__ void __launch_bounds__(kBlock) k_flush_scatter(unsigned long long *acc, double *rho, double charge,
                                                         size_t n, PhBuild ph, unsigned ntab, PushArgs a,
                                                         unsigned long long *next_len)
{
    const unsigned ns = MCC ? kScatterBlocks : 0u;
    if (MCC && blockIdx.x < ns) {
        scatter_queue(a, (unsigned long long)blockIdx.x * kBlock + threadIdx.x,
                      (unsigned long long)kScatterBlocks * kBlock, next_len);
        return;
    }
    if (blockIdx.x < ns + ntab) {
        const size_t base = (size_t)(blockIdx.x - ns) * (kBlock * kPhPerThread) + threadIdx.x;
#pragma unroll
        for (int r = 0; r < kPhPerThread; r++) {
            const size_t e = base + (size_t)r * kBlock;
            if (e < ph.nxe) {
                const float2 g = ph.gx[e];
                ph.phx[e] = fmaf(ph.ph, g.y, g.x);
            } else if (e < ph.nxe + ph.nze) {
                const float2 g = ph.gz[e - ph.nxe];
                ph.phz[e - ph.nxe] = fmaf(ph.ph, g.y, g.x);
            }
        }
        return;
    }
    const size_t k = (size_t)(blockIdx.x - ns - ntab) * kBlock + threadIdx.x;
    if (k >= n) return;
    const long long v = (long long)acc[k];
    if (v) {
        rho[k] += charge * ((double)v * kInvFix);
        acc[k] = 0ull;
    }
}

// Species<CARTESIAN>::advance_boris_init (particles.cpp:997-1050)
__global__ void __launch_bounds__(kBlock) k_advance_init2d(PushArgs a)
{
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= a.n) return;
    float x = a.x[p];
    if (is_dead(x)) return;
    float z = a.z[p];
    float vx = a.vx[p], vy = a.vy[p], vz = a.vz[p];
    float ex, ez;
    if (a.xl) gather_E(a, pos64(x, a.xl[p]), pos64(z, a.zl[p]), a.phase, ex, ez);
    else gather_E(a, (double)x, (double)z, a.phase, ex, ez);
    float tx = a.tx, ty = a.ty, tz = a.tz, sx = a.sx, sy = a.sy, sz = a.sz;
    if (a.bt && local_rotation(a, x, z, tx, ty, tz, sx, sy, sz)) raise_flag(a.counters, M2D_FLAG_OUT_OF_RANGE);
    float vpr = vx - vy * tz + vz * ty;
    float vpz = vz - vx * ty + vy * tx;
    float vpt = vy - vz * tx + vx * tz;
    vx = vx - vpt * sz + vpz * sy;
    vy = vy - vpz * sx + vpr * sz;
    vz = vz - vpr * sy + vpt * sx;
    vx = fmaf(ex, a.qmdt2, vx);
    vz = fmaf(ez, a.qmdt2, vz);
    a.vx[p] = vx; a.vy[p] = vy; a.vz[p] = vz;
}

// Fields::B at FP64 points (magnetic_field_const = 0: the B grid's tables)
__global__ void k_fields_B(PushArgs a, const double *xs, const double *zs, int64_t n, double *br, double *bz)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    float r, z;
    if (gather_B(a, xs[k], zs[k], r, z)) raise_flag(a.counters, M2D_FLAG_OUT_OF_RANGE);
    br[k] = r;
    bz[k] = z;
}

__global__ void k_fill_f64(double *d, double v, int64_t n)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) d[k] = v;
}

static int fill_f64(m2d_ctx *ctx, double *d, double v, size_t n)
{
    k_fill_f64<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(d, v, (int64_t)n);
    M2D_HIP(hipGetLastError());
    return M2D_OK;
}

// Fields::E at FP64 points through the same tables (test entry point)
__global__ void k_fields_E(PushArgs a, const double *xs, const double *zs, int64_t n, double *ex,
                           double *ez)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    float fx, fz;
    gather_E(a, xs[k], zs[k], a.phase, fx, fz);
    ex[k] = fx;
    ez[k] = fz;
}

// Species::accumulate (particles.hpp:358-367) into the fixed-point grid
__global__ void __launch_bounds__(kBlock) k_accumulate2d(const float *xs, const float *zs, const float *xls,
                                                         const float *zls, int64_t n, int nx, int nz, double idx,
                                                         double idz, double xmin, double zmin,
                                                         unsigned long long *acc,
                                                         unsigned long long *counters)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float x = xs[p];
    if (is_dead(x)) return;
    const float xl = xls ? xls[p] : 0.f, zl = zls ? zls[p] : 0.f;
    int i, j;
    uint32_t w[4];
    bool in_grid;
    if (xmin == 0.0 && zmin == 0.0) {
        // the FP32 pair arithmetic of the specialised pass, so a grid of
        // origin 0 gets the bits of the fused deposit (a NaN fails x >= 0)
        const float z = zs[p];
        float sxh, sxl, szh, szl, u, ou, v, ov;
        split_f32(idx, sxh, sxl);
        split_f32(idz, szh, szl);
        in_grid = pos64(x, xl) >= 0.0 && pos64(z, zl) >= 0.0;
        if (in_grid) {
            cell_f(x, xl, sxh, sxl, i, u, ou);
            cell_f(z, zl, szh, szl, j, v, ov);
            in_grid = (unsigned)i <= (unsigned)(nx - 2) && (unsigned)j <= (unsigned)(nz - 2);
            if (in_grid) cic_from_offsets(u, ou, v, ov, w);
        }
    } else {
        in_grid = cic_weights((pos64(x, xl) - xmin) * idx, (pos64(zs[p], zl) - zmin) * idz, nx, nz, i, j, w);
    }
    if (in_grid) {
        const long long ww[4] = {w[0], w[1], w[2], w[3]};
        deposit_global(acc, nz, i, j, ww);
    }
    else
        raise_flag(counters, M2D_FLAG_OUT_OF_RANGE);
}

// rho += charge * acc * 2^-31; acc = 0
__global__ void k_flush(unsigned long long *acc, double *rho, double charge, size_t n)
{
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    long long v = (long long)acc[k];
    if (v) {
        rho[k] += charge * ((double)v * kInvFix);
        acc[k] = 0ull;
    }
}

// Field2D::interpolate (Field2D.hpp:64-78) in FP64
__global__ void k_interp2d(const double *d, int nx, int nz, double idx, double idz, double xmin,
                           double zmin, const double *xs, const double *zs, int64_t n, double *out,
                           unsigned long long *counters)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double x = xs[k] - xmin, z = zs[k] - zmin;
    if (!(x >= 0) || !(z >= 0)) { out[k] = 0; raise_flag(counters, M2D_FLAG_OUT_OF_RANGE); return; }
    int i = (int)(x * idx), j = (int)(z * idz);
    if (i > nx - 2 || j > nz - 2) { out[k] = 0; raise_flag(counters, M2D_FLAG_OUT_OF_RANGE); return; }
    double u = x * idx - i, v = z * idz - j;
    const double *r = d + (size_t)i * nz + j;
    out[k] = (1 - u) * (1 - v) * r[0] + u * (1 - v) * r[nz] + (1 - u) * v * r[1] + u * v * r[nz + 1];
}

// Field2D::grad (Field2D.hpp:80-168) in FP64, branch for branch
__global__ void k_grad2d(const double *d, int nx, int nz, double idx, double idy, double xmin,
                         double zmin, const double *xs, const double *zs, int64_t n, double *gx,
                         double *gz)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int jmax = nx, lmax = nz;
    auto D = [&](int a, int b) { return d[(size_t)a * nz + b]; };
    double x = xs[k] - xmin, y = zs[k] - zmin;
    double g1, g2, g3, g4, fx, fy;
    int i = (int)(x * idx + 0.5), j = (int)(y * idy);
    j = min(j, lmax - 2);
    if (j >= 0) {
        if (i > 0 && i < jmax - 1) {
            g1 = (D(i, j) - D(i - 1, j)) * idx;
            g2 = (D(i, j + 1) - D(i - 1, j + 1)) * idx;
            g3 = (D(i + 1, j + 1) - D(i, j + 1)) * idx;
            g4 = (D(i + 1, j) - D(i, j)) * idx;
            fx = x * idx - i + .5;
            fy = y * idy - j;
            gx[k] = g1 * (1 - fx) * (1 - fy) + g2 * (1 - fx) * fy + g3 * fx * fy + g4 * fx * (1 - fy);
        } else if (i == jmax - 1) {
            g1 = (D(i, j) - D(i - 1, j)) * idx;
            g2 = (D(i, j + 1) - D(i - 1, j + 1)) * idx;
            fy = y * idy - j;
            gx[k] = g1 * (1 - fy) + g2 * fy;
        } else if (i == 0) {
            g3 = (D(i + 1, j + 1) - D(i, j + 1)) * idx;
            g4 = (D(i + 1, j) - D(i, j)) * idx;
            fy = y * idy - j;
            gx[k] = g3 * fy + g4 * (1 - fy);
        }
    }
    i = (int)(x * idx);
    j = (int)(y * idy + 0.5);
    i = min(i, jmax - 2);
    if (i >= 0) {
        if (j > 0 && j < lmax - 1) {
            g1 = (D(i, j) - D(i, j - 1)) * idy;
            g2 = (D(i + 1, j) - D(i + 1, j - 1)) * idy;
            g3 = (D(i + 1, j + 1) - D(i + 1, j)) * idy;
            g4 = (D(i, j + 1) - D(i, j)) * idy;
            fx = x * idx - i;
            fy = y * idy - j + 0.5;
            gz[k] = g1 * (1 - fx) * (1 - fy) + g2 * (1 - fy) * fx + g3 * fx * fy + g4 * fy * (1 - fx);
        } else if (j == lmax - 1) {
            g1 = (D(i, j) - D(i, j - 1)) * idy;
            g2 = (D(i + 1, j) - D(i + 1, j - 1)) * idy;
            fx = x * idx - i;
            gz[k] = g1 * (1 - fx) + g2 * fx;
        } else if (j == 0) {
            g3 = (D(i + 1, j + 1) - D(i + 1, j)) * idy;
            g4 = (D(i, j + 1) - D(i, j)) * idy;
            fx = x * idx - i;
            gz[k] = g3 * fx + g4 * (1 - fx);
        }
    }
}

// ------------------------------------------------------------------ reductions
// per-block sum of |v|^2 (FP64) and live count; fixed-order tree in LDS
__global__ void __launch_bounds__(kBlock) k_energy_partial(const float *xs, const float *vx,
                                                           const float *vy, const float *vz,
                                                           int64_t n, double *psum,
                                                           unsigned long long *pcnt)
{
    __shared__ double s[kBlock];
    __shared__ unsigned long long c[kBlock];
    double acc = 0;
    unsigned long long cnt = 0;
    const int64_t base = (int64_t)blockIdx.x * (kBlock * kPPT) + threadIdx.x;
    for (int q = 0; q < kPPT; q++) {
        int64_t p = base + (int64_t)q * kBlock;
        if (p < n && !is_dead(xs[p])) {
            double a = vx[p], b = vy[p], d = vz[p];
            acc += a * a + d * d + b * b;
            cnt++;
        }
    }
    s[threadIdx.x] = acc;
    c[threadIdx.x] = cnt;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            s[threadIdx.x] += s[threadIdx.x + o];
            c[threadIdx.x] += c[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { psum[blockIdx.x] = s[0]; pcnt[blockIdx.x] = c[0]; }
}

__global__ void __launch_bounds__(1024) k_energy_final(const double *psum,
                                                       const unsigned long long *pcnt, int nb,
                                                       double *out_sum,
                                                       unsigned long long *out_cnt)
{
    __shared__ double s[1024];
    __shared__ unsigned long long c[1024];
    double a = 0;
    unsigned long long b = 0;
    for (int k = threadIdx.x; k < nb; k += 1024) { a += psum[k]; b += pcnt[k]; }
    s[threadIdx.x] = a;
    c[threadIdx.x] = b;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) { s[threadIdx.x] += s[threadIdx.x + o]; c[threadIdx.x] += c[threadIdx.x + o]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { *out_sum = s[0]; *out_cnt = c[0]; }
}

// ------------------------------------------------------------------ compaction
constexpr int kCompactItems = kBlock * kPPT;

__global__ void __launch_bounds__(kBlock) k_count_blocks(const float *xs, int64_t n,
                                                         unsigned int *counts)
{
    __shared__ unsigned int s[kBlock / 64];
    unsigned int c = 0;
    const int64_t base = (int64_t)blockIdx.x * kCompactItems + threadIdx.x;
    for (int q = 0; q < kPPT; q++) {
        int64_t p = base + (int64_t)q * kBlock;
        c += (p < n && !is_dead(xs[p])) ? 1u : 0u;
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if (__lane_id() == 0) s[threadIdx.x / 64] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int t = 0;
        for (int w = 0; w < kBlock / 64; w++) t += s[w];
        counts[blockIdx.x] = t;
    }
}

// exclusive scan of nb block counts by one block; total into *total
__global__ void __launch_bounds__(1024) k_scan_counts(const unsigned int *counts, int nb,
                                                      unsigned long long *offs,
                                                      unsigned long long *total)
{
    __shared__ unsigned long long s[1024];
    const int per = (nb + 1023) / 1024;
    const int lo = threadIdx.x * per, hi = min(nb, lo + per);
    unsigned long long loc = 0;
    for (int k = lo; k < hi; k++) loc += counts[k];
    s[threadIdx.x] = loc;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        unsigned long long v = (int)threadIdx.x >= o ? s[threadIdx.x - o] : 0ull;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    unsigned long long run = s[threadIdx.x] - loc;
    for (int k = lo; k < hi; k++) { offs[k] = run; run += counts[k]; }
    if (threadIdx.x == 1023) *total = s[1023];
}

__global__ void __launch_bounds__(kBlock) k_compact_scatter(m2d_particles2d src, m2d_particles2d dst,
                                                            const unsigned long long *offs)
{
    __shared__ unsigned int wsum[kBlock / 64];
    unsigned long long run = offs[blockIdx.x];
    const int lane = __lane_id(), wid = threadIdx.x / 64;
    const int64_t base = (int64_t)blockIdx.x * kCompactItems + threadIdx.x;
    for (int q = 0; q < kPPT; q++) {
        int64_t p = base + (int64_t)q * kBlock;
        bool live = p < src.n && !is_dead(src.x[p]);
        unsigned long long m = __ballot(live);
        unsigned int before = (unsigned int)__popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wid] = (unsigned int)__popcll(m);
        __syncthreads();
        unsigned int woff = 0, tot = 0;
        for (int w = 0; w < kBlock / 64; w++) {
            woff += w < wid ? wsum[w] : 0u;
            tot += wsum[w];
        }
        if (live) {
            unsigned long long d = run + woff + before;
            dst.x[d] = src.x[p]; dst.z[d] = src.z[p];
            dst.vx[d] = src.vx[p]; dst.vy[d] = src.vy[p]; dst.vz[d] = src.vz[p];
            if (src.xl) { dst.xl[d] = src.xl[p]; dst.zl[d] = src.zl[p]; }
        }
        run += tot;
        __syncthreads();
    }
}

__global__ void k_fill_dead(float *x, int64_t n)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) x[k] = dead_value();
}

// ------------------------------------------------------------------ sampler
// a sampled FP64 position: rounded to FP32, with its remainder in a store
// that keeps long-time positions
__device__ __forceinline__ void store_pos(const m2d_particles2d &P, int64_t p, double x, double z)
{
    if (P.xl) {
        pos_split(x, P.x[p], P.xl[p]);
        pos_split(z, P.z[p], P.zl[p]);
    } else {
        P.x[p] = (float)x;
        P.z[p] = (float)z;
    }
}

// add_particles_on_disk (particles.cpp:728-749), one Philox stream per slot
__global__ void __launch_bounds__(kBlock) k_disk2d(m2d_particles2d P, double cx, double cz,
                                                   double R, double vth, double bx0, double bx1,
                                                   double bz0, double bz1, uint64_t seed,
                                                   uint32_t stream)
{
    int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= P.n) return;
    Rng r;
    r.begin(seed, stream, (uint32_t)p, 0, kPurposeInit);
    double x = 1, y = 1;
    for (int t = 0; t < 64 && x * x + y * y > 0.25; t++) {
        x = r.uni() - 0.5;
        y = r.uni() - 0.5;
    }
    x = x * 2 * R + cx;
    y = y * 2 * R + cz;
    if (x < bx0 || x > bx1 || y < bz0 || y > bz1) {
        P.x[p] = dead_value();
        return;
    }
    double a = r.normal() * vth, b = r.normal() * vth, c = r.normal() * vth;
    store_pos(P, p, x, y);
    P.vx[p] = (float)a; P.vz[p] = (float)b; P.vy[p] = (float)c;
}

// J0 by its power series sum_k (-x^2/4)^k / (k!)^2; for 0 <= x <= 2.5 the
// terms fall below 1e-20 within 20 terms (the sampler below needs x <= 2.41)
__device__ __forceinline__ double bessel_j0_series(double x)
{
    const double q = -0.25 * x * x;
    double t = 1.0, s = 1.0;
#pragma unroll
    for (int k = 1; k <= 20; k++) {
        t *= q / (double)(k * k);
        s += t;
    }
    return s;
}

// add_particles_bessel (particles.cpp:701-726): density J0(2.405 r / R) on the
// disk by rejection from the square, one Philox stream per slot
__global__ void __launch_bounds__(kBlock) k_bessel2d(m2d_particles2d P, double cx, double cz, double R,
                                                     double vth, double bx0, double bx1, double bz0, double bz1,
                                                     uint64_t seed, uint32_t stream)
{
    int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= P.n) return;
    Rng r;
    r.begin(seed, stream, (uint32_t)p, 0, kPurposeInit);
    const double root = 2.404825557695773;  // first zero of J0 (particles.cpp:708)
    double x = 0, y = 0;
    // acceptance is about 1/3 per trial: 256 trials fail with probability 1e-46
    for (int t = 0; t < 256; t++) {
        x = r.uni() * 2 - 1.0;
        y = r.uni() * 2 - 1.0;
        const double rr = sqrt(x * x + y * y);
        if (rr > 1.0) continue;
        if (!(r.uni() > bessel_j0_series(rr * root))) break;
    }
    x = x * R + cx;
    y = y * R + cz;
    if (x < bx0 || x > bx1 || y < bz0 || y > bz1) {
        P.x[p] = dead_value();
        return;
    }
    // rndv(vx, vz, vy) (particles.hpp:188-193)
    double a = r.normal() * vth, b = r.normal() * vth, c = r.normal() * vth;
    store_pos(P, p, x, y);
    P.vx[p] = (float)a; P.vz[p] = (float)b; P.vy[p] = (float)c;
}

// add_particles_everywhere (particles.cpp:686-699): uniform in the box, one
// Philox stream per slot; time_to_death (may be null) = lifetime Exp(1)
__global__ void __launch_bounds__(kBlock) k_everywhere2d(m2d_particles2d P, double x0, double x1, double z0,
                                                         double z1, double vth, double *ttd, double lifetime,
                                                         uint64_t seed, uint32_t stream)
{
    int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= P.n) return;
    Rng r;
    r.begin(seed, stream, (uint32_t)p, 0, kPurposeInit);
    const double x = (x1 - x0) * r.uni() + x0;
    (void)r.uni();  // y: the reference draws it; the Cartesian store has no y
    const double z = (z1 - z0) * r.uni() + z0;
    const double a = r.normal() * vth, b = r.normal() * vth, c = r.normal() * vth;
    store_pos(P, p, x, z);
    P.vx[p] = (float)a; P.vy[p] = (float)b; P.vz[p] = (float)c;
    if (ttd) ttd[p] = lifetime * r.exponential();
}

// ------------------------------------------------------------------ host helpers
int check_parts(const m2d_particles2d *p)
{
    M2D_CHECK(p, M2D_ERR_BAD_ARG, "particles is NULL");
    M2D_CHECK(p->n >= 0 && p->n < (1ll << 32), M2D_ERR_BAD_ARG, "particle count %lld out of range",
              (long long)p->n);
    M2D_CHECK(p->n == 0 || (p->x && p->z && p->vx && p->vy && p->vz), M2D_ERR_BAD_ARG,
              "particle arrays must be non-NULL");
    uintptr_t al = (uintptr_t)p->x | (uintptr_t)p->z | (uintptr_t)p->vx | (uintptr_t)p->vy |
                   (uintptr_t)p->vz;
    M2D_CHECK((al & 15) == 0, M2D_ERR_BAD_ARG, "particle arrays must be 16-byte aligned");
    M2D_CHECK((p->xl == nullptr) == (p->zl == nullptr), M2D_ERR_BAD_ARG,
              "position remainders xl and zl must both be set or both be NULL");
    M2D_CHECK((((uintptr_t)p->xl | (uintptr_t)p->zl) & 15) == 0, M2D_ERR_BAD_ARG,
              "position remainder arrays must be 16-byte aligned");
    return M2D_OK;
}

int ensure_acc(m2d_ctx *ctx, size_t n)
{
    size_t need = n * sizeof(unsigned long long);
    if (ctx->acc.bytes >= need) return M2D_OK;
    M2D_TRY(ensure(ctx, ctx->acc, need));
    return zero_words(ctx, (unsigned long long *)ctx->acc.ptr, ctx->acc.bytes / 8);
}

static int flush(m2d_ctx *ctx, size_t n, double *rho, double charge)
{
    k_flush<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>((unsigned long long *)ctx->acc.ptr,
                                                                   rho, charge, n);
    M2D_HIP(hipGetLastError());
    return M2D_OK;
}

// the smallest float >= v and the largest float <= v: for a float x,
// x >= v iff x >= f32_at_least(v), and x <= v iff x <= f32_at_most(v)
static float f32_at_least(double v)
{
    float f = (float)v;
    if ((double)f < v) f = nextafterf(f, INFINITY);
    return f;
}
static float f32_at_most(double v)
{
    float f = (float)v;
    if ((double)f > v) f = nextafterf(f, -INFINITY);
    return f;
}

static void fill_push_args(PushArgs &a, m2d_fields2d *f, const m2d_species2d *s,
                           const m2d_particles2d *p, double rot_scale, double kick_scale)
{
    const m2d_fields2d_params &P = f->p;
    a.x = p->x; a.z = p->z; a.vx = p->vx; a.vy = p->vy; a.vz = p->vz; a.n = p->n;
    a.xl = p->xl; a.zl = p->zl;
    a.gx = f->gx; a.gz = f->gz; a.cellfree = f->cellfree; a.winfree = f->winfree;
    a.phx = f->phx; a.phz = f->phz;
    a.nx = f->nx; a.nz = f->nz; a.idx = f->idx; a.idz = f->idz;
    split_f32(f->idx, a.idxh, a.idxl);
    split_f32(f->idz, a.idzh, a.idzl);
    a.xmin = P.grid.xmin; a.zmin = P.grid.zmin;
    a.fx0 = f32_at_least(P.grid.xmin); a.fx1 = f32_at_most(P.x_max);
    a.fz0 = f32_at_least(P.grid.zmin); a.fz1 = f32_at_most(P.z_max);
    a.xper = P.x_max; a.zper = P.z_max;
    // rotation: t = B q dt/(2m) scaled (particles.cpp:959-975, init: -0.5, 1028)
    double tmp = rot_scale * s->charge * s->dt / (2.0 * s->mass);
    double tx = P.Br * tmp, ty = P.Bt * tmp, tz = P.Bz * tmp;
    double sc = 2.0 / (1 + tx * tx + ty * ty + tz * tz);
    a.tx = (float)tx; a.ty = (float)ty; a.tz = (float)tz;
    a.sx = (float)(tx * sc); a.sy = (float)(ty * sc); a.sz = (float)(tz * sc);
    a.rotate = tx != 0.0 || ty != 0.0 || tz != 0.0;
    a.push = 1;
    a.qmdt2 = (float)(kick_scale * s->charge / s->mass * s->dt / 2.0);
    a.dt = (float)s->dt;
    a.dtl = (float)(s->dt - (double)a.dt);
    // Fields::E (fields.hpp:124-150); phase hoisted: t = niter dt for every particle
    a.econst = P.geometry_empty && !P.selfconsistent;
    a.ecx = 0.f;
    a.ecz = (float)P.extern_field;
    double phase = 0.0;
    if (P.rf) phase = P.rf_amplitude * rf_cos(s->niter, s->dt, P.rf_omega) + P.rf_U0;
    a.phase = (float)phase;
    a.boundary = 0; a.periodic = P.boundary == M2D_BC_PERIODIC;
    a.checkfree = !P.electric_field_from_file;
    a.trial_thr = 0; a.seed = f->ctx->seed; a.stream = s->stream; a.step = s->niter;
    a.mcc = nullptr;
    a.cq = nullptr;
    a.cqv = nullptr;
    a.cq_cap = 0;
    a.cq_defer = 0;
    a.cq_len = nullptr;
    a.sched = nullptr;
    a.n_chunks = nullptr;
    a.ntz = 0;
    a.acc = nullptr;
    a.counters = f->ctx->counters;
    a.bt = f->bt;
    a.bnx = a.bnz = 0;
    a.bidx = a.bidz = a.bxmin = a.bzmin = 0.0;
    a.rscale = tmp;
    if (f->bt) {
        // Fields::B with magnetic_field_const = 0: the params' B is not used
        a.bnx = f->bgrid.nx; a.bnz = f->bgrid.nz;
        a.bidx = 1.0 / f->bgrid.dx; a.bidz = 1.0 / f->bgrid.dz;
        a.bxmin = f->bgrid.xmin; a.bzmin = f->bgrid.zmin;
        a.tx = a.ty = a.tz = a.sx = a.sy = a.sz = 0.f;
        a.rotate = 1;
    }
}

// The compiled form of the pass for one call's configuration: the store is
// sorted (tiled), the configuration is the specialised one (fast), the
// phase-applied tables hold this step's phase (pht), B comes from a B grid,
// the deposit and the MCC trial are on.  Long-time positions (HIPOS) have no
// generic tiled form: the caller runs those untiled.
template <bool HIPOS>
static void (*pick_pass(bool tiled, bool fast, bool pht, bool bgrid, bool deposit, bool mcc))(PushArgs)
{
    if (tiled) {
        if (fast && pht)
            return mcc ? k_advance2d<true, true, true, true, false, true, HIPOS>
                       : k_advance2d<true, false, true, true, false, true, HIPOS>;
        if (fast)
            return mcc ? k_advance2d<true, true, true, true, false, false, HIPOS>
                       : k_advance2d<true, false, true, true, false, false, HIPOS>;
        if constexpr (HIPOS) {
            return nullptr;
        } else {
            if (bgrid) return mcc ? k_advance2d<true, true, true, false, true> : k_advance2d<true, false, true, false, true>;
            return mcc ? k_advance2d<true, true, true, false> : k_advance2d<true, false, true, false>;
        }
    }
    if (fast) {
        if (deposit) return mcc ? k_advance2d<true, true, false, true, false, false, HIPOS>
                                : k_advance2d<true, false, false, true, false, false, HIPOS>;
        return mcc ? k_advance2d<false, true, false, true, false, false, HIPOS>
                   : k_advance2d<false, false, false, true, false, false, HIPOS>;
    }
    if (bgrid) {
        if (deposit) return mcc ? k_advance2d<true, true, false, false, true, false, HIPOS>
                                : k_advance2d<true, false, false, false, true, false, HIPOS>;
        return mcc ? k_advance2d<false, true, false, false, true, false, HIPOS>
                   : k_advance2d<false, false, false, false, true, false, HIPOS>;
    }
    if (deposit) return mcc ? k_advance2d<true, true, false, false, false, false, HIPOS>
                            : k_advance2d<true, false, false, false, false, false, HIPOS>;
    return mcc ? k_advance2d<false, true, false, false, false, false, HIPOS>
               : k_advance2d<false, false, false, false, false, false, HIPOS>;
}

}  // namespace m2d

using namespace m2d;

extern "C" {

int m2d_advance_2d(m2d_ctx *ctx, m2d_fields2d *f, const m2d_species2d *s, const m2d_particles2d *p,
                   const m2d_mcc *mcc, double *rho, uint32_t flags, m2d_step_stats *stats)
{
    M2D_CHECK(ctx && f && s, M2D_ERR_BAD_ARG, "m2d_advance_2d: NULL argument");
    M2D_CHECK(f->ctx == ctx, M2D_ERR_BAD_ARG, "m2d_advance_2d: fields belong to another context");
    M2D_TRY(check_parts(p));
    M2D_CHECK(s->mass > 0 && s->dt > 0, M2D_ERR_BAD_ARG, "m2d_advance_2d: mass and dt must be > 0");
    bool deposit = (flags & M2D_STEP_DEPOSIT) != 0;
    M2D_CHECK(!deposit || rho, M2D_ERR_BAD_ARG, "m2d_advance_2d: DEPOSIT needs rho");
    M2D_CHECK((flags & ~(M2D_STEP_PUSH | M2D_STEP_BOUNDARY | M2D_STEP_DEPOSIT)) == 0, M2D_ERR_BAD_ARG,
              "m2d_advance_2d: unknown flags 0x%x", flags);
    if (stats) {
        M2D_TRY(zero_words(ctx, ctx->counters + CNT_REMOVED, 2));
    }
    PushArgs a;
    fill_push_args(a, f, s, p, 1.0, 1.0);
    a.boundary = (flags & M2D_STEP_BOUNDARY) != 0;
    a.push = (flags & M2D_STEP_PUSH) != 0;
    // the MCC trial belongs to advance_boris; an empty store queues nothing
    // (and leaves the alternating queue counters as they are)
    bool use_mcc = p->n > 0 && a.push && mcc && std::isfinite(s->lifetime) && s->lifetime > 0;
    unsigned long long *next_len = nullptr;
    if (use_mcc) {
        a.trial_thr = trial_threshold(1.0 - exp(-s->dt / s->lifetime));  // particles.cpp:937
        a.mcc = mcc->d;
        // slots (u32), then velocities (float4) from a 16-byte boundary
        const size_t qn = (size_t)p->n, vo = (qn * sizeof(uint32_t) + 15) & ~(size_t)15;
        M2D_TRY(ensure(ctx, ctx->cq, vo + qn * sizeof(float4)));
        a.cq_cap = (uint32_t)std::min<size_t>(qn, 0xFFFFFFFFu);
        a.cq = (uint32_t *)ctx->cq.ptr;
        a.cqv = (float4 *)((uint8_t *)ctx->cq.ptr + vo);
        // two queue counters alternate between launches: a launch's scatter
        // kernel clears the one the next launch appends to.  A call that
        // failed between its pass and its scatter left them unknown: clear
        // both first.
        if (ctx->cq_dirty) M2D_TRY(zero_words(ctx, ctx->counters + CNT_QLEN, 2));
        ctx->cq_dirty = true;  // until this call's scatter is queued
        const int slot = (int)(ctx->launch_seq++ & 1);
        a.cq_len = ctx->counters + CNT_QLEN + slot;
        next_len = ctx->counters + CNT_QLEN + (slot ^ 1);
    }
    // a store sorted by m2d_sort_2d against these fields deposits through
    // per-chunk LDS windows
    const m2d_tiling *T = deposit ? find_tiling(ctx, p->x, p->n) : nullptr;
    if (T && (T->nty != 1 || T->ntx != (f->nx - 2) / kTile + 1 || T->ntz != (f->nz - 2) / kTile + 1)) T = nullptr;
    // the common configuration runs the compile-time specialised pass
    const bool fast = a.push && a.boundary && a.checkfree && !a.rotate && !a.periodic && !a.econst &&
                      a.xmin == 0.0 && a.zmin == 0.0;
    // long-time positions: the generic forms are built untiled only
    const bool hipos = p->xl != nullptr;
    if (hipos && !fast) T = nullptr;
    if (deposit) {
        M2D_TRY(ensure_acc(ctx, f->n));
        a.acc = (unsigned long long *)ctx->acc.ptr;
    }
    if (T) {
        a.sched = T->sched;
        a.n_chunks = T->n_chunks;
        a.ntz = T->ntz;
    }
    bool fast_tiled = false;  // this call ran the specialised tiled pass
    if (p->n > 0) {
        const bool bgrid = f->bt != nullptr;  // B per particle (generic pass)
        // the flush after the previous pass left the tables for this phase
        const bool pht = T && fast && f->ph_valid && f->ph_phase == a.phase;
        void (*kern)(PushArgs) = hipos ? pick_pass<true>(T != nullptr, fast, pht, bgrid, deposit, use_mcc)
                                       : pick_pass<false>(T != nullptr, fast, pht, bgrid, deposit, use_mcc);
        M2D_CHECK(kern, M2D_ERR_UNSUPPORTED, "m2d_advance_2d: no compiled form of the pass for this configuration");
        const int g = T ? (int)xcd_run_grid((unsigned)T->max_chunks) : (int)((p->n + kChunk - 1) / kChunk);
        fast_tiled = T && fast;
        launch_timed(ctx, kern, dim3(g), dim3(kPass), a);
        M2D_HIP(hipGetLastError());
        ctx->last_pass = (deposit ? M2D_PASS_DEPOSIT : 0u) | (use_mcc ? M2D_PASS_MCC : 0u) |
                         (T ? M2D_PASS_TILED : 0u) | (fast ? M2D_PASS_FAST : bgrid ? M2D_PASS_BGRID : 0u) |
                         (pht ? M2D_PASS_PHT : 0u) | (hipos ? M2D_PASS_HIPOS : 0u);
    }
    if (deposit || use_mcc) {
        // the grid flush, the next step's phase-applied tables (after a
        // specialised tiled pass, when they do not hold that phase yet) and
        // the collision queue in one launch.  The next phase is the one the
        // next call computes when it steps niter by one; any other call finds
        // the tables stale and gathers from gx, gz.
        const unsigned nfl = deposit ? (unsigned)((f->n + kBlock - 1) / kBlock) : 0u;
        double phn = 0.0;
        if (f->p.rf) phn = f->p.rf_amplitude * rf_cos(s->niter + 1, s->dt, f->p.rf_omega) + f->p.rf_U0;
        PhBuild pb;
        pb.gx = f->gx; pb.gz = f->gz; pb.phx = f->phx; pb.phz = f->phz;
        pb.nxe = (size_t)(f->nx + 1) * f->nz;
        pb.nze = (size_t)f->nx * (f->nz + 1);
        pb.ph = (float)phn;
        // (self-consistent fields are solved again before the next step, which
        // rebuilds gx, gz and drops tables built now)
        const bool build_ph = fast_tiled && !f->p.selfconsistent && !(f->ph_valid && f->ph_phase == pb.ph);
        const unsigned ntab =
            build_ph ? (unsigned)((pb.nxe + pb.nze + kBlock * kPhPerThread - 1) / (kBlock * kPhPerThread)) : 0u;
        // an interaction that turns its partner (COULOMB) scatters the queue
        // in slot order after the flush, by one thread
        const bool serial = use_mcc && mcc->h.partner_update;
        // a target whose partner ensemble is this very store would read
        // velocities that other collisions of the step are writing: the
        // scatter then keeps its results in the queue, and k_scatter_apply
        // stores them once every collision has read its partner
        bool self_target = false;
        if (use_mcc && !serial)
            for (int t = 0; t < mcc->h.n_targets; t++) self_target |= mcc->h.t[t].px == p->x;
        a.cq_defer = self_target ? 1 : 0;
        const unsigned g = nfl + ntab + (use_mcc && !serial ? kScatterBlocks : 0u);
        // without a deposit no flush block runs: n = 0 ends them at once
        const size_t nodes = deposit ? f->n : 0;
        if (use_mcc && !serial)
            k_flush_scatter<true><<<g, kBlock, 0, ctx->stream>>>(a.acc, rho, s->charge, nodes, pb, ntab, a, next_len);
        else if (g > 0)
            k_flush_scatter<false><<<g, kBlock, 0, ctx->stream>>>(a.acc, rho, s->charge, nodes, pb, ntab, a,
                                                                  next_len);
        M2D_HIP(hipGetLastError());
        if (a.cq_defer) {
            k_scatter_apply<<<kScatterBlocks, kBlock, 0, ctx->stream>>>(a);
            M2D_HIP(hipGetLastError());
        }
        if (build_ph) {
            f->ph_valid = true;
            f->ph_phase = pb.ph;
        }
        if (serial) {
            unsigned long long len = 0;
            M2D_HIP(hipMemcpyAsync(&len, a.cq_len, sizeof(len), hipMemcpyDeviceToHost, ctx->stream));
            M2D_HIP(hipStreamSynchronize(ctx->stream));
            uint32_t *sorted = nullptr;
            if (len > 0) {
                M2D_TRY(ensure(ctx, ctx->tmp, (size_t)len * sizeof(uint32_t) + 256));
                sorted = (uint32_t *)ctx->tmp.ptr;
                M2D_TRY(sort_keys_u32(ctx, a.cq, sorted, (size_t)len));
            }
            k_scatter_serial<<<1, 64, 0, ctx->stream>>>(a, sorted, len, next_len);
            M2D_HIP(hipGetLastError());
        }
    }
    if (use_mcc) ctx->cq_dirty = false;
    if (stats) {
        M2D_TRY(read_counters(ctx));
        stats->removed = (int64_t)ctx->h_counters[CNT_REMOVED];
        stats->collisions = (int64_t)ctx->h_counters[CNT_COLL];
        stats->spilled = 0;  // not counted by the 2D pass
    }
    return M2D_OK;
}

int m2d_advance_init_2d(m2d_ctx *ctx, m2d_fields2d *f, const m2d_species2d *s,
                        const m2d_particles2d *p)
{
    M2D_CHECK(ctx && f && s, M2D_ERR_BAD_ARG, "m2d_advance_init_2d: NULL argument");
    M2D_TRY(check_parts(p));
    M2D_CHECK(s->mass > 0 && s->dt > 0, M2D_ERR_BAD_ARG, "m2d_advance_init_2d: mass and dt must be > 0");
    PushArgs a;
    fill_push_args(a, f, s, p, -0.5, -1.0);
    if (p->n > 0) {
        k_advance_init2d<<<grid_for(p->n, 1), kBlock, 0, ctx->stream>>>(a);
        M2D_HIP(hipGetLastError());
    }
    return M2D_OK;
}

int m2d_fields2d_E(m2d_fields2d *f, const double *x, const double *z, int64_t n, double t,
                   double *ex, double *ez)
{
    M2D_CHECK(f && (n == 0 || (x && z && ex && ez)), M2D_ERR_BAD_ARG, "m2d_fields2d_E: NULL argument");
    m2d_species2d s{};
    s.charge = 1; s.mass = 1; s.dt = 1;
    // phase at time t: niter*dt == t with dt = t, niter = 1 (t > 0) or dt = 1, niter = 0
    if (t > 0) { s.dt = t; s.niter = 1; }
    PushArgs a;
    m2d_particles2d none{};
    fill_push_args(a, f, &s, &none, 1.0, 1.0);
    if (n > 0) {
        k_fields_E<<<(unsigned)((n + 255) / 256), 256, 0, f->ctx->stream>>>(a, x, z, n, ex, ez);
        M2D_HIP(hipGetLastError());
    }
    M2D_HIP(hipStreamSynchronize(f->ctx->stream));
    return M2D_OK;
}

int m2d_fields2d_B(m2d_fields2d *f, const double *x, const double *z, int64_t n, double *Br, double *Bz,
                   double *Bt)
{
    M2D_CHECK(f && (n == 0 || (x && z && Br && Bz && Bt)), M2D_ERR_BAD_ARG, "m2d_fields2d_B: NULL argument");
    if (n == 0) return M2D_OK;
    if (!f->bt) {
        // magnetic_field_const (fields.hpp:154-169)
        M2D_TRY(fill_f64(f->ctx, Br, f->p.Br, (size_t)n));
        M2D_TRY(fill_f64(f->ctx, Bz, f->p.Bz, (size_t)n));
        M2D_TRY(fill_f64(f->ctx, Bt, f->p.Bt, (size_t)n));
    } else {
        m2d_species2d s{};
        s.charge = 1; s.mass = 1; s.dt = 1;
        PushArgs a;
        m2d_particles2d none{};
        fill_push_args(a, f, &s, &none, 1.0, 1.0);
        k_fields_B<<<(unsigned)((n + 255) / 256), 256, 0, f->ctx->stream>>>(a, x, z, n, Br, Bz);
        M2D_HIP(hipGetLastError());
        M2D_TRY(fill_f64(f->ctx, Bt, 0.0, (size_t)n));
    }
    M2D_HIP(hipStreamSynchronize(f->ctx->stream));
    return M2D_OK;
}

int m2d_accumulate_2d(m2d_ctx *ctx, const m2d_grid2d *g, const m2d_particles2d *p, double charge,
                      double *rho)
{
    M2D_CHECK(ctx && g && rho, M2D_ERR_BAD_ARG, "m2d_accumulate_2d: NULL argument");
    M2D_CHECK(g->nx >= 2 && g->nz >= 2 && g->dx > 0 && g->dz > 0, M2D_ERR_BAD_ARG,
              "m2d_accumulate_2d: bad grid");
    M2D_TRY(check_parts(p));
    size_t n = (size_t)g->nx * g->nz;
    M2D_TRY(ensure_acc(ctx, n));
    if (p->n > 0) {
        k_accumulate2d<<<grid_for(p->n, 1), kBlock, 0, ctx->stream>>>(
            p->x, p->z, p->xl, p->zl, p->n, g->nx, g->nz, 1.0 / g->dx, 1.0 / g->dz, g->xmin, g->zmin,
            (unsigned long long *)ctx->acc.ptr, ctx->counters);
        M2D_HIP(hipGetLastError());
    }
    return flush(ctx, n, rho, charge);
}

int m2d_interpolate_2d(m2d_ctx *ctx, const m2d_grid2d *g, const double *data, const double *x,
                       const double *z, int64_t n, double *out)
{
    M2D_CHECK(ctx && g && data && (n == 0 || (x && z && out)), M2D_ERR_BAD_ARG,
              "m2d_interpolate_2d: NULL argument");
    if (n > 0) {
        k_interp2d<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(
            data, g->nx, g->nz, 1.0 / g->dx, 1.0 / g->dz, g->xmin, g->zmin, x, z, n, out,
            ctx->counters);
        M2D_HIP(hipGetLastError());
    }
    return M2D_OK;
}

int m2d_grad_2d(m2d_ctx *ctx, const m2d_grid2d *g, const double *data, const double *x,
                const double *z, int64_t n, double *gx, double *gz)
{
    M2D_CHECK(ctx && g && data && (n == 0 || (x && z && gx && gz)), M2D_ERR_BAD_ARG,
              "m2d_grad_2d: NULL argument");
    if (n > 0) {
        k_grad2d<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(
            data, g->nx, g->nz, 1.0 / g->dx, 1.0 / g->dz, g->xmin, g->zmin, x, z, n, gx, gz);
        M2D_HIP(hipGetLastError());
    }
    return M2D_OK;
}

int m2d_mean_energy_2d(m2d_ctx *ctx, const m2d_particles2d *p, double mass, double *mean_eV,
                       int64_t *n_live)
{
    M2D_CHECK(ctx && mean_eV, M2D_ERR_BAD_ARG, "m2d_mean_energy_2d: NULL argument");
    M2D_TRY(check_parts(p));
    int nb = grid_for(p->n, kPPT);
    M2D_TRY(ensure(ctx, ctx->red, (size_t)nb * 16 + 64));
    double *psum = (double *)ctx->red.ptr;
    unsigned long long *pcnt = (unsigned long long *)(psum + nb);
    double *osum = (double *)(pcnt + nb);
    unsigned long long *ocnt = (unsigned long long *)(osum + 1);
    k_energy_partial<<<nb, kBlock, 0, ctx->stream>>>(p->x, p->vx, p->vy, p->vz, p->n, psum, pcnt);
    k_energy_final<<<1, 1024, 0, ctx->stream>>>(psum, pcnt, nb, osum, ocnt);
    M2D_HIP(hipGetLastError());
    double hs = 0;
    unsigned long long hc = 0;
    M2D_HIP(hipMemcpyAsync(&hs, osum, 8, hipMemcpyDeviceToHost, ctx->stream));
    M2D_HIP(hipMemcpyAsync(&hc, ocnt, 8, hipMemcpyDeviceToHost, ctx->stream));
    M2D_HIP(hipStreamSynchronize(ctx->stream));
    // Histogram::mean_tot = sum_tot / n_val_tot (histogram.cpp:68-71)
    *mean_eV = hc ? hs * mass * 0.5 / kQE / (double)hc : 0.0;
    if (n_live) *n_live = (int64_t)hc;
    return M2D_OK;
}

int m2d_count_live_2d(m2d_ctx *ctx, const m2d_particles2d *p, int64_t *n_live)
{
    M2D_CHECK(ctx && n_live, M2D_ERR_BAD_ARG, "m2d_count_live_2d: NULL argument");
    M2D_TRY(check_parts(p));
    int nb = grid_for(p->n, kPPT);
    M2D_TRY(ensure(ctx, ctx->tmp, (size_t)nb * 12 + 64));
    unsigned int *counts = (unsigned int *)ctx->tmp.ptr;
    unsigned long long *offs = (unsigned long long *)(((uintptr_t)(counts + nb) + 15) & ~(uintptr_t)15);
    unsigned long long *total = offs + nb;
    k_count_blocks<<<nb, kBlock, 0, ctx->stream>>>(p->x, p->n, counts);
    k_scan_counts<<<1, 1024, 0, ctx->stream>>>(counts, nb, offs, total);
    M2D_HIP(hipGetLastError());
    unsigned long long h = 0;
    M2D_HIP(hipMemcpyAsync(&h, total, 8, hipMemcpyDeviceToHost, ctx->stream));
    M2D_HIP(hipStreamSynchronize(ctx->stream));
    *n_live = (int64_t)h;
    return M2D_OK;
}

int m2d_compact_2d(m2d_ctx *ctx, const m2d_particles2d *src, const m2d_particles2d *dst,
                   int64_t *n_live)
{
    M2D_CHECK(ctx && n_live, M2D_ERR_BAD_ARG, "m2d_compact_2d: NULL argument");
    M2D_TRY(check_parts(src));
    M2D_TRY(check_parts(dst));
    M2D_CHECK(dst->n >= src->n, M2D_ERR_BAD_ARG, "m2d_compact_2d: dst holds %lld < %lld slots",
              (long long)dst->n, (long long)src->n);
    M2D_CHECK(dst->x != src->x, M2D_ERR_BAD_ARG, "m2d_compact_2d: src and dst must differ");
    M2D_CHECK((src->xl == nullptr) == (dst->xl == nullptr), M2D_ERR_BAD_ARG,
              "m2d_compact_2d: src and dst must both hold position remainders or neither");
    int nb = grid_for(src->n, kPPT);
    M2D_TRY(ensure(ctx, ctx->tmp, (size_t)nb * 12 + 64));
    unsigned int *counts = (unsigned int *)ctx->tmp.ptr;
    unsigned long long *offs = (unsigned long long *)(((uintptr_t)(counts + nb) + 15) & ~(uintptr_t)15);
    unsigned long long *total = offs + nb;
    k_count_blocks<<<nb, kBlock, 0, ctx->stream>>>(src->x, src->n, counts);
    k_scan_counts<<<1, 1024, 0, ctx->stream>>>(counts, nb, offs, total);
    forget_tiling(ctx, dst->x);
    k_compact_scatter<<<nb, kBlock, 0, ctx->stream>>>(*src, *dst, offs);
    M2D_HIP(hipGetLastError());
    unsigned long long h = 0;
    M2D_HIP(hipMemcpyAsync(&h, total, 8, hipMemcpyDeviceToHost, ctx->stream));
    M2D_HIP(hipStreamSynchronize(ctx->stream));
    *n_live = (int64_t)h;
    // slots past the live prefix become empty
    if ((int64_t)h < dst->n) {
        int64_t rest = dst->n - (int64_t)h;
        k_fill_dead<<<(unsigned)((rest + 255) / 256), 256, 0, ctx->stream>>>(dst->x + h, rest);
        M2D_HIP(hipGetLastError());
    }
    return M2D_OK;
}

int m2d_add_particles_on_disk_2d(m2d_ctx *ctx, m2d_fields2d *f, const m2d_particles2d *p, double cx,
                                 double cz, double radius, double vth, uint32_t stream)
{
    M2D_CHECK(ctx && f, M2D_ERR_BAD_ARG, "m2d_add_particles_on_disk_2d: NULL argument");
    M2D_TRY(check_parts(p));
    M2D_CHECK(radius > 0 && vth >= 0, M2D_ERR_BAD_ARG, "m2d_add_particles_on_disk_2d: radius/vth");
    if (p->n == 0) return M2D_OK;
    forget_tiling(ctx, p->x);
    k_disk2d<<<grid_for(p->n, 1), kBlock, 0, ctx->stream>>>(*p, cx, cz, radius, vth, f->p.grid.xmin,
                                                             f->p.x_max, f->p.grid.zmin, f->p.z_max,
                                                             ctx->seed, stream);
    M2D_HIP(hipGetLastError());
    return M2D_OK;
}

int m2d_add_particles_bessel_2d(m2d_ctx *ctx, m2d_fields2d *f, const m2d_particles2d *p, double cx, double cz,
                                double radius, double vth, uint32_t stream)
{
    M2D_CHECK(ctx && f, M2D_ERR_BAD_ARG, "m2d_add_particles_bessel_2d: NULL argument");
    M2D_TRY(check_parts(p));
    M2D_CHECK(radius > 0 && vth >= 0, M2D_ERR_BAD_ARG, "m2d_add_particles_bessel_2d: radius/vth");
    if (p->n == 0) return M2D_OK;
    forget_tiling(ctx, p->x);
    k_bessel2d<<<grid_for(p->n, 1), kBlock, 0, ctx->stream>>>(*p, cx, cz, radius, vth, f->p.grid.xmin,
                                                               f->p.x_max, f->p.grid.zmin, f->p.z_max,
                                                               ctx->seed, stream);
    M2D_HIP(hipGetLastError());
    return M2D_OK;
}

int m2d_add_particles_everywhere_2d(m2d_ctx *ctx, m2d_fields2d *f, const m2d_particles2d *p, double vth,
                                    double *time_to_death, double lifetime, uint32_t stream)
{
    M2D_CHECK(ctx && f, M2D_ERR_BAD_ARG, "m2d_add_particles_everywhere_2d: NULL argument");
    M2D_TRY(check_parts(p));
    M2D_CHECK(vth >= 0, M2D_ERR_BAD_ARG, "m2d_add_particles_everywhere_2d: vth must be >= 0");
    M2D_CHECK(!time_to_death || lifetime > 0, M2D_ERR_BAD_ARG,
              "m2d_add_particles_everywhere_2d: lifetime must be > 0");
    if (p->n == 0) return M2D_OK;
    forget_tiling(ctx, p->x);
    k_everywhere2d<<<grid_for(p->n, 1), kBlock, 0, ctx->stream>>>(*p, f->p.grid.xmin, f->p.x_max,
                                                                   f->p.grid.zmin, f->p.z_max, vth,
                                                                   time_to_death, lifetime, ctx->seed, stream);
    M2D_HIP(hipGetLastError());
    return M2D_OK;
}

}  // extern "C"
