// m2d_internal.hpp -- shared internals of libmag2d_amd (HIP, gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <cstdint>
#include <cmath>
#include <cstdio>
#include <cstdarg>
#include <string>
#include <vector>
#include "mag2d_amd.h"

namespace m2d {

// ---------------------------------------------------------------- constants
// physical constants of the reference (param.cpp:7-11)
constexpr double kEps0 = 8.854187817e-12;
constexpr double kKB = 1.380662e-23;
constexpr double kQE = 1.602189e-19;

constexpr double kInvFix = 1.0 / 2147483648.0;  // 2^-31: unit of the fixed-point deposit grid
constexpr int kBlock = 256;                  // threads per block of the particle kernels

// Philox counter word 3 carries (purpose << 24)
constexpr uint32_t kPurposeMcc = 0;    // scatter draws of one slot: counter (slot, step, block)
constexpr uint32_t kPurposeInit = 1;   // initial sampling of one slot
constexpr uint32_t kPurposeTrial = 2;  // MCC trial: counter (slot/4, step, 0), word slot%4
constexpr uint32_t kPurposeMulticoll = 3;  // multicoll events of one slot in one step: the
                                           // time_to_death draws and scatter, in event order
constexpr uint32_t kPurposeTtd = 4;    // the first time_to_death of one slot

constexpr int kTile = 8;               // cells per side of the sort tile
constexpr int kTileShift = 6;          // key = tile << 6 | cell in tile
constexpr int kHalo = 8;               // deposit window halo (cells) around a tile: a particle
                                       // stays in its window for many steps after a sort
constexpr int kWin = kTile + 1 + 2 * kHalo;  // window side in nodes
constexpr int kChunkCap = 1984;        // particles per block of the tiled 2D pass: with its
                                       // start rounded down to 64 slots, two staged rounds of 512
constexpr int kRound2 = 1024;          // slots per round of the tiled 2D pass (256 threads x 4)
constexpr int kChunkCap3 = 1984;       // particles per block of the tiled 3D pass

// ---------------------------------------------------------------- errors
int set_error(int code, const char *fmt, ...);
#define M2D_HIP(call)                                                                 \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess)                                                         \
            return ::m2d::set_error(M2D_ERR_HIP, "%s: %s (%s:%d)", #call,             \
                                    hipGetErrorString(e_), __FILE__, __LINE__);       \
    } while (0)
#define M2D_CHECK(cond, code, ...)                                                    \
    do {                                                                              \
        if (!(cond)) return ::m2d::set_error(code, __VA_ARGS__);                      \
    } while (0)
#define M2D_TRY(call)                                                                 \
    do {                                                                              \
        int r_ = (call);                                                              \
        if (r_ != M2D_OK) return r_;                                                  \
    } while (0)

// ---------------------------------------------------------------- XCD order
// The dispatcher deals a launch's blocks round-robin to the chip's 8 XCDs
// (block b to XCD b % 8), each with its own L2.  xcd_block turns that into
// contiguous ranges: the blocks of XCD x take logical indices
// [x G, (x + 1) G), G = ceil(n / 8), in dispatch order, so neighbouring
// logical blocks (neighbouring tiles, or destination ranges of a nearly
// sorted gather) share one L2.  A launch needs at least 8 G blocks (n
// rounded up to a multiple of 8); indices >= n come back for the blocks
// with nothing to do.
constexpr unsigned kXcds = 8;
inline unsigned xcd_grid(unsigned n) { return (n + kXcds - 1) / kXcds * kXcds; }
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned n)
{
    const unsigned G = (n + kXcds - 1) / kXcds;
    return b < kXcds * G ? (b % kXcds) * G + b / kXcds : 0xFFFFFFFFu;
}
// The same in runs of kXcdRun consecutive logical blocks dealt round-robin
// to the XCDs: XCD x takes runs x, x + 8, x + 16, ... in order.  Neighbours within
// a run share an L2, and work that varies smoothly along the logical order
// (dense and sparse regions of a sorted store) spreads over every XCD.  A
// launch needs xcd_run_grid(n) blocks; indices >= n mean nothing to do.
constexpr unsigned kXcdRun = 8;
inline unsigned xcd_run_grid(unsigned n)
{
    return (n + kXcds * kXcdRun - 1) / (kXcds * kXcdRun) * (kXcds * kXcdRun);
}
__device__ __forceinline__ unsigned xcd_run_block(unsigned b)
{
    const unsigned x = b % kXcds, j = b / kXcds;
    return ((j / kXcdRun) * kXcds + x) * kXcdRun + j % kXcdRun;
}

// ---------------------------------------------------------------- device buffers
// A growable device scratch buffer owned by a context.
struct Scratch {
    void *ptr = nullptr;
    size_t bytes = 0;
};

}  // namespace m2d

// Tile schedule of a tile-sorted store (built by m2d_sort_2d): the live
// particles of sort tile t are slots [tile_start[t], tile_end[t]); chunk c
// of the tiled pass covers sched[c] = (tile, start, end), at most kChunkCap
// (2D) or kChunkCap3 (3D) particles of one tile.  Valid for the store at `x` with `n` slots until
// the store is written by anything but the particle pass.
struct m2d_tiling {
    const float *x = nullptr;
    int64_t n = 0;
    int ntx = 0, nty = 1, ntz = 0;  // sort tiles along x, (y,) z
    int max_chunks = 0;        // launch bound
    m2d::Scratch buf;          // device: sched, tile_start, tile_end, chunk_count, n_chunks
    int *tile_start = nullptr, *tile_end = nullptr, *chunk_count = nullptr;
    int4 *sched = nullptr;
    unsigned int *n_chunks = nullptr;
};

// counters written by kernels
// CNT_SPILL: deposits of a tiled pass that fell outside their tile's LDS
// window (summed straight into the grid)
enum { CNT_REMOVED = 0, CNT_COLL = 1, CNT_FLAGS = 2, CNT_LIVE = 3, CNT_QLEN = 4 /* and 5 */, CNT_SPILL = 6, CNT_N = 8 };

struct m2d_ctx {
    int refs = 1;           // the creator's reference plus one per Fields / MCC object
    int device = 0;
    uint64_t seed = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    unsigned long long *counters = nullptr;  // CNT_N device words
    unsigned long long *h_counters = nullptr; // pinned mirror
    m2d::Scratch acc;       // int64 fixed-point deposit grid
    m2d::Scratch red;       // reduction partials
    m2d::Scratch tmp;       // general temporaries (compaction offsets)
    m2d::Scratch sortbuf;   // keys, slots and radix-sort temporaries
    m2d::Scratch cq;        // MCC collision queue (slot indices)
    uint64_t launch_seq = 0;  // selects the collision-queue counter of a launch
    bool cq_dirty = false;    // a call failed between appending to a queue counter and its scatter
    uint32_t last_pass = 0;   // M2D_PASS_* bits of the last 2D particle pass launched
    std::vector<struct m2d_tiling *> tilings;  // schedules of recently sorted stores
    // kernel timing: event pairs around particle-pass launches
    bool profile = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pending;
};

namespace m2d {
// tile schedules of sorted stores (m2d_sort.hip)
m2d_tiling *find_tiling(m2d_ctx *ctx, const float *x, int64_t n);
int sort_keys_u32(m2d_ctx *ctx, const uint32_t *in, uint32_t *out, size_t n);
int build_tiling(m2d_ctx *ctx, const float *x, int64_t n, const uint32_t *keys, uint32_t dead_key, int ntx,
                 int nty, int ntz, int cap, int round);
void forget_tiling(m2d_ctx *ctx, const float *x);
void free_tilings(m2d_ctx *ctx);
// objects that hold a context keep it alive until they are destroyed
void ctx_retain(m2d_ctx *ctx);
void ctx_release(m2d_ctx *ctx);
// bracket a launch with events when profiling is on
// Kernel timing (m2d_ctx_profile): launch_timed launches through
// hipExtLaunchKernel with start and stop events that the runtime stamps at the
// kernel's own dispatch and completion, so the pair holds the kernel alone
// (an event pair recorded around a launch can also hold the wait of an idle
// queue for the launch).
bool prof_pair(m2d_ctx *ctx, hipEvent_t &start, hipEvent_t &stop);

template <typename... Args, typename F = void (*)(Args...)>
inline void launch_timed(m2d_ctx *ctx, F kernel, dim3 grid, dim3 block, Args... args)
{
    hipEvent_t e0, e1;
    if (prof_pair(ctx, e0, e1)) hipExtLaunchKernelGGL(kernel, grid, block, 0, ctx->stream, e0, e1, 0, args...);
    else hipLaunchKernelGGL(kernel, grid, block, 0, ctx->stream, args...);
}
}  // namespace m2d

namespace m2d {

int ensure(m2d_ctx *ctx, Scratch &s, size_t bytes);

// vector kernels shared by the 2D and 3D solvers (m2d_fields2d.hip) and the
// fixed-point flush (m2d_particles2d.hip)
// The solver kernels take `stop`, a device flag that the solve's own control
// kernels set once an iteration is not to run: a launch whose flag is set
// returns at once (the flag is set before the launch starts and read the same
// by every thread, so a whole launch runs or none of it).  Null: always run.
__device__ __forceinline__ bool stopped(const int *stop) { return stop && *stop; }
// the CG direction at node k: p' = z + beta p (first: p' = z), beta = rz_new / rz_old
__device__ __forceinline__ float cg_beta(const double *rz_new, const double *rz_old)
{
    const double o = *rz_old;
    return o != 0.0 ? (float)(*rz_new / o) : 0.f;
}
__device__ __forceinline__ float cg_dir_at(const float *z, const float *p, size_t k, float beta, int first)
{
    return first ? z[k] : fmaf(beta, p[k], z[k]);
}
__global__ void k_flush(unsigned long long *acc, double *rho, double charge, size_t n);
// Species::accumulate into a fixed-point grid (m2d_particles2d.hip)
__global__ void __launch_bounds__(kBlock) k_accumulate2d(const float *xs, const float *zs, const float *xls,
                                                         const float *zls, int64_t n, int nx, int nz, double idx,
                                                         double idz, double xmin, double zmin,
                                                         unsigned long long *acc, unsigned long long *counters);
// device address of the no-allocation entry point's static scratch on the
// current device, resolved once per device (m2d_noalloc.hip)
int noalloc_init_device(int device);
__global__ void k_dot_fin(const float *a, const float *b, size_t n, double *partial, unsigned int *ticket,
                          double *out, int fin, const int *stop);
constexpr int kCoarseThreads = 1024;
constexpr int kDenseParts = 4;  // threads per row of the dense coarsest solves
__global__ void k_dense_solve(const float *invT, const float *b, float *x, int n, const int *stop);
int ensure_acc(m2d_ctx *ctx, size_t n);
int fill_bytes(m2d_ctx *ctx, void *dst, int value, size_t bytes);
int zero_words(m2d_ctx *ctx, unsigned long long *dst, size_t n);
int read_counters(m2d_ctx *ctx);

// ---------------------------------------------------------------- solve control
// The 2D and 3D solves' control block, in device memory after the 16 scalars
// of their dot array ([0], [4] r.z of alternate CG iterations, [1] p.q, [2]
// r.r, [8] |r|^2 and [9] |b_eff|^2 of the FP64 residual, [10] |d|^2 of its
// FP32 copy).  Small kernels of one thread keep it between the launches, so
// the refinement and its CG passes are queued whole and read on the host
// once: a pass that has converged turns the launches queued after it into
// launches that return at once (see stopped()).
constexpr int kMaxOuter = 12;            // refinement passes at most
constexpr int kMaxPasses = kMaxOuter + 1;
struct SolveCtl {
    int stop_outer;      // refinement done (converged, failed or out of passes): nothing more runs
    int stop_pcg;        // no CG pass open, or the open one has converged
    int no_add;          // the pass's correction is not to be added
    int outer;           // corrections added
    int iters;           // CG iterations run in all passes
    int pass_iters;      // in the open pass
    int max_pass_iters;  // the most in one pass
    int nan;             // the residual was not finite
    int halted;          // a pass ran out of queued iterations: suspended until the host queues more
    double rel;          // FP64 relative residual at the last evaluation
    double thr;          // the open pass's stop: |r| <= rtol |d|
    int pass_count[kMaxPasses];  // CG iterations of each pass
};
static_assert(sizeof(SolveCtl) <= 16 * sizeof(double), "SolveCtl fits its 16 words");

// the FP32 correction is solved to 1e-5 of its right-hand side, or only as
// far as the FP64 target needs (half of tol / rel, at most 0.1): the last
// refinement pass then takes one or two CG iterations, not five
__host__ __device__ inline double inner_rtol(double rel, double tol)
{
    return fmax(1e-5, fmin(0.1, 0.5 * tol / rel));
}


// The control updates that read a reduction's sums run in the thread that
// writes them, at the end of the reduction (finish_sums / k_sum_partials):
// - after the FP64 residual's norms (|r|^2, |b_eff|^2): the relative
//   residual and whether the refinement is done;
// - after a CG iteration's |r|^2: the iteration counts, and whether the
//   open pass has converged.
// Both reductions are launches gated by the flag they may set, so a stopped
// solve skips them with the rest.
struct CtlEval {
    SolveCtl *c;
    double tol;
    int max_outer;
    __device__ void operator()(const double *out) const
    {
        const double rn = sqrt(out[0]), bn = sqrt(out[1]);
        const double rel = bn > 0 ? rn / bn : (rn > 0 ? INFINITY : 0.0);
        c->rel = rel;
        if (!isfinite(rn)) {
            c->nan = 1;
            c->stop_outer = 1;
        } else if (rel <= tol || c->outer >= max_outer) {
            c->stop_outer = 1;
        }
    }
};
struct CtlCheck {
    SolveCtl *c;
    __device__ void operator()(const double *out) const
    {
        c->iters++;
        c->pass_iters++;
        if (c->pass_iters > c->max_pass_iters) c->max_pass_iters = c->pass_iters;
        if (c->outer < kMaxPasses) c->pass_count[c->outer] = c->pass_iters;
        if (sqrt(out[0]) <= c->thr) c->stop_pcg = 1;
    }
};
// a reduction with no follow-up
struct NoPost {
    __device__ void operator()(const double *) const {}
};
// e += alpha p, r -= alpha q and |r|^2, then post (m2d_fields2d.hip)
__global__ void k_cg_update_dot(float *e, float *r, const float *p, const float *q, const double *rz,
                                const double *pq, size_t n, int first, double *partial, unsigned int *ticket,
                                double *out, int fin, const int *stop, CtlCheck post);

// the last solve's shape on a set of fields: refinement passes, and the CG
// iterations each pass took (the next solve queues as many, plus one)
struct SolveEst {
    int passes = 2, max_iters = 6;
    int iters[kMaxPasses] = {6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6};
};
static_assert(kMaxPasses == 13, "SolveEst::iters initialiser");

__global__ void k_ctl_init(SolveCtl *c);
__global__ void k_ctl_open(SolveCtl *c, const double *dots, double tol);
__global__ void k_ctl_end(SolveCtl *c, int max_iters);
__global__ void k_ctl_resume(SolveCtl *c);
__global__ void k_ctl_close(SolveCtl *c);

// FP64 iterative refinement over FP32 CG passes, queued under SolveCtl.  The
// caller has queued the right-hand side, k_ctl_init and resid() (the FP64
// residual, its norms and CtlEval).  Passes are queued as many as the
// last solve took, each with one CG iteration more than its counterpart
// took then (the last pass of a refinement needs far fewer than the first);
// the control block is read once after them.  A pass that needs more
// iterations than queued is suspended on the device (k_ctl_end) and resumed
// in the next round with twice as many; a refinement that needs more passes
// gets one more a round.  So the iterations that run, and the result, are
// those of a loop that tests on the host after each, whatever the estimates.
// pcg(from, to, cur) queues iterations from + 1 .. to of the open pass,
// add() the correction's addition (which skips itself while ctl->no_add).
template <class Pcg, class Add, class Resid>
int queue_refinement(hipStream_t s, SolveCtl *ctl, SolveCtl *h, const double *dots, SolveEst &est, double tol,
                     int max_iters, Pcg pcg, Add add, Resid resid)
{
    auto finish_pass = [&] {
        k_ctl_end<<<1, 64, 0, s>>>(ctl, max_iters);
        add();
        k_ctl_close<<<1, 64, 0, s>>>(ctl);
        resid();
    };
    int J = est.passes < 1 ? 1 : (est.passes > kMaxOuter ? kMaxOuter : est.passes);
    int queued = 0;      // passes opened so far
    int resume = 0;      // iterations the suspended pass has run (0: none suspended)
    // each round opens at least one pass or lengthens a suspended one by at
    // least 4 of its max_iters iterations: a bound the rounds cannot reach
    const int max_rounds = (kMaxOuter + 1) * (max_iters / 4 + 2);
    for (int round = 0; round < max_rounds; round++) {
        if (resume) {
            int cur = (resume - 1) & 1;
            const int to = resume + (resume > 4 ? resume : 4);
            k_ctl_resume<<<1, 64, 0, s>>>(ctl);
            if (int r = pcg(resume, to < max_iters ? to : max_iters, cur)) return r;
            finish_pass();
        }
        for (int j = 0; j < J; j++, queued++) {
            const int e = queued < kMaxPasses ? est.iters[queued] : est.max_iters;
            int K = e + 1;
            K = K < 2 ? 2 : (K > max_iters ? max_iters : K);
            int cur = 0;
            k_ctl_open<<<1, 64, 0, s>>>(ctl, dots, tol);
            if (int r = pcg(0, K, cur)) return r;
            finish_pass();
        }
        M2D_HIP(hipGetLastError());
        M2D_HIP(hipMemcpyAsync(h, ctl, sizeof(SolveCtl), hipMemcpyDeviceToHost, s));
        M2D_HIP(hipStreamSynchronize(s));
        if (h->halted) {
            // pass h->outer stopped after all its queued iterations: passes
            // queued after it never opened
            resume = h->pass_iters;
            queued = h->outer + 1;
            J = est.passes - queued > 0 ? est.passes - queued : 0;
            continue;
        }
        if (h->nan || h->stop_outer) break;
        resume = 0;
        J = 1;
    }
    est.passes = h->outer > 1 ? h->outer : 1;
    est.max_iters = h->max_pass_iters;
    for (int j = 0; j < kMaxPasses; j++) est.iters[j] = j < h->outer ? h->pass_count[j] : h->max_pass_iters;
    return M2D_OK;
}

// null, count, non-null and 16-byte aligned arrays of a 2D store
int check_parts(const m2d_particles2d *p);

// ---------------------------------------------------------------- device helpers
__device__ __forceinline__ bool is_dead(float x) { return __float_as_uint(x) == M2D_DEAD_BITS; }
__device__ __forceinline__ float dead_value() { return __uint_as_float(M2D_DEAD_BITS); }

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
__device__ __forceinline__ void philox4x32_10(uint32_t &c0, uint32_t &c1, uint32_t &c2,
                                              uint32_t &c3, uint32_t k0, uint32_t k1)
{
    // each 32 x 32 -> 64 product is one v_mad_u64_u32 (the separate low and
    // high multiplies would cost two)
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0; c1 = (uint32_t)p1; c2 = n2; c3 = (uint32_t)p0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
}

// Draw stream of one (slot, step, purpose): words come in blocks of four.
struct Rng {
    uint32_t k0, k1, slot, step_lo, hi3, block;
    uint32_t b0, b1, b2, b3;
    int used;  // words consumed from the current block

    __device__ __forceinline__ void begin(uint64_t seed, uint32_t stream, uint32_t s,
                                          uint64_t step, uint32_t purpose)
    {
        k0 = (uint32_t)seed;
        k1 = (uint32_t)(seed >> 32) ^ stream;
        slot = s;
        step_lo = (uint32_t)step;
        hi3 = (purpose << 24) | ((uint32_t)(step >> 32) & 0xFFFFFFu);
        block = 0;
        used = 4;
    }
    __device__ __forceinline__ void refill()
    {
        b0 = slot; b1 = step_lo; b2 = block; b3 = hi3;
        philox4x32_10(b0, b1, b2, b3, k0, k1);
        block++;
        used = 0;
    }
    __device__ __forceinline__ uint32_t u32()
    {
        if (used == 4) refill();
        uint32_t w = used == 0 ? b0 : used == 1 ? b1 : used == 2 ? b2 : b3;
        used++;
        return w;
    }
    // uniform in (0,1), exact in FP64: (w + 0.5) 2^-32
    __device__ __forceinline__ double uni() { return ((double)u32() + 0.5) * 2.3283064365386963e-10; }
    // Box-Muller, cosine branch (two uniforms per variate)
    __device__ __forceinline__ double normal()
    {
        double a = uni(), b = uni();
        return sqrt(-2.0 * log(a)) * cos(2.0 * M_PI * b);
    }
    // Exp(1) by inversion; t_random::rexp (random.cpp:54) draws the same
    // distribution by the ziggurat
    __device__ __forceinline__ double exponential() { return -log(uni()); }
};

// The MCC trial draw is uni() = (w + 1/2) 2^-32 for a 32-bit word w; since
// scaling by 2^32 is exact, uni() < prob iff w < ceil(prob 2^32 - 1/2).
inline uint64_t trial_threshold(double prob)
{
    if (!(prob > 0.0)) return 0;
    if (prob >= 1.0) return 1ull << 32;
    return (uint64_t)ceil(prob * 4294967296.0 - 0.5);
}

// The trial word of slot p at a step: slots 4m..4m+3 share the Philox block
// (counter (m, step, 0, purpose 2), key (seed, seed >> 32 ^ stream)), word p & 3
__device__ __forceinline__ uint32_t trial_word(uint64_t seed, uint32_t stream, uint64_t p, uint64_t step)
{
    uint32_t c0 = (uint32_t)(p >> 2), c1 = (uint32_t)step, c2 = 0u;
    uint32_t c3 = (kPurposeTrial << 24) | ((uint32_t)(step >> 32) & 0xFFFFFFu);
    philox4x32_10(c0, c1, c2, c3, (uint32_t)seed, (uint32_t)(seed >> 32) ^ stream);
    const int w = (int)(p & 3);
    return w == 0 ? c0 : w == 1 ? c1 : w == 2 ? c2 : c3;
}

// ---------------------------------------------------------------- long-time positions
// A store with position remainders (m2d_particles2d::xl, zl) holds each
// position as an FP32 pair whose exact FP64 sum is the position.
__device__ __forceinline__ double pos64(float hi, float lo) { return (double)hi + (double)lo; }
__device__ __forceinline__ void pos_split(double v, float &hi, float &lo)
{
    hi = (float)v;
    lo = (float)(v - (double)hi);
}
// hi + lo += v dt with dt = dth + dtl an FP32 pair: the product v dth is
// d plus its rounding error (an fma), plus v dtl; Knuth's two-sum of hi and
// d is exact, and its error joins the product's error and lo in the new
// remainder; a fast two-sum renormalises the pair.  The pair then carries
// the position to about 2^-48 of its size, where FP32 alone keeps 2^-24.
__device__ __forceinline__ void drift_pair(float &hi, float &lo, float v, float dth, float dtl)
{
    const float d = v * dth;
    const float dl = fmaf(v, dth, -d) + v * dtl;
    const float s = hi + d;
    const float bp = s - hi;
    const float e = (hi - (s - bp)) + (d - bp);
    const float l = lo + (e + dl);
    hi = s + l;
    lo = l - (hi - s);
}

// wave-aggregated counter increment (one atomic per wavefront)
__device__ __forceinline__ void wave_count(unsigned long long *ctr, bool pred)
{
    unsigned long long m = __ballot(pred);
    if (m) {
        int lane = __lane_id();
        int first = __ffsll((long long)m) - 1;
        if (lane == first) atomicAdd(ctr, (unsigned long long)__popcll(m));
    }
}

__device__ __forceinline__ void raise_flag(unsigned long long *counters, unsigned long long f)
{
    atomicOr(&counters[CNT_FLAGS], f);
}

// cos of the RF phase at step niter: t = niter dt, cos(mod(omega t, 1e7 pi))
// (Fields::E, fields.hpp:140-144; mod, mymath.cpp:62-67), on the host
inline double rf_cos(uint64_t niter, double dt, double omega)
{
    const double t = (double)niter * dt;
    double ph = t * omega;
    const double y = 10000000 * M_PI;
    if (!(ph >= 0.0 && ph <= y)) ph = ph - y * (int)(ph / y) + (ph < 0 ? y : 0);
    return cos(ph);
}

// v as an FP32 pair: hi = v rounded, lo = the rounded remainder
__host__ __device__ inline void split_f32(double v, float &hi, float &lo)
{
    hi = (float)v;
    lo = (float)(v - (double)hi);
}

inline int grid_for(int64_t n, int per_thread, int block = kBlock)
{
    int64_t g = (n + (int64_t)block * per_thread - 1) / ((int64_t)block * per_thread);
    return (int)(g < 1 ? 1 : g);
}


// Reductions of the solver kernels (m2d_fields2d.hip, m2d_3d.hip).  Each
// block sums its values in a fixed order and writes a partial.  The partials
// are then summed in block order, by sum_partials, either in a kernel of its
// own (k_sum_partials) or, with fin set, by the block that takes the last
// ticket, which writes out[] and resets the ticket: no extra launch, which
// pays on small grids.  The last-ticket path costs each block a device-scope
// fence, so it is for launches of at most kFinBlocks blocks.  Both paths add
// in the same order and give the same bits on every run.  Blocks of
// kRedBlock threads; every thread of every block must call finish_sums.
constexpr int kRedBlock = 256;
constexpr unsigned kFinBlocks = 256;

template <int NV>
__device__ __forceinline__ void sum_partials(const double *partial, unsigned nb, double *out)
{
    __shared__ double sh[NV][kRedBlock / 64];
    const int w = threadIdx.x >> 6;
    double t[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) t[q] = 0;
    // unrolled so that several loads are in flight; the adds keep block order
#pragma unroll 8
    for (unsigned b = threadIdx.x; b < nb; b += kRedBlock)
#pragma unroll
        for (int q = 0; q < NV; q++) t[q] += partial[NV * b + q];
#pragma unroll
    for (int q = 0; q < NV; q++) {
        for (int o = 32; o > 0; o >>= 1) t[q] += __shfl_down(t[q], o);
        if ((threadIdx.x & 63) == 0) sh[q][w] = t[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) {
            double s = 0;
            for (int k = 0; k < kRedBlock / 64; k++) s += sh[q][k];
            out[q] = s;
        }
    }
}

// post(out): what the thread that wrote the sums does with them next
template <int NV, class Post = NoPost>
__global__ void __launch_bounds__(kRedBlock) k_sum_partials(const double *partial, unsigned nb, double *out,
                                                           const int *stop, Post post = Post{})
{
    if (stopped(stop)) return;
    sum_partials<NV>(partial, nb, out);
    if (threadIdx.x == 0) post(out);
}

template <int NV, class Post = NoPost>
__device__ __forceinline__ void finish_sums(double (&v)[NV], double *partial, unsigned int *ticket, double *out,
                                            int fin, Post post = Post{})
{
    __shared__ double sh[NV][kRedBlock / 64];
    __shared__ bool last;
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; q++) {
        for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_down(v[q], o);
        if ((threadIdx.x & 63) == 0) sh[q][w] = v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) {
            double t = 0;
            for (int k = 0; k < kRedBlock / 64; k++) t += sh[q][k];
            partial[NV * blockIdx.x + q] = t;
        }
        if (fin) {
            __threadfence();
            last = atomicAdd(ticket, 1u) == gridDim.x - 1;
        }
    }
    if (!fin) return;  // uniform: a kernel argument
    __syncthreads();
    if (!last) return;  // uniform over the block
    __threadfence();
    sum_partials<NV>(partial, gridDim.x, out);
    if (threadIdx.x == 0) {
        *ticket = 0u;
        post(out);
    }
}

// launch a reduction kernel over nb blocks and finish its sums: in the
// kernel (fin) when nb <= kFinBlocks, else with k_sum_partials (the kernel
// then leaves post to it: pass the same post to both)
template <int NV, class Launch, class Post = NoPost>
inline void reduce_launch(unsigned nb, double *partial, double *out, hipStream_t s, Launch launch,
                          const int *stop = nullptr, Post post = Post{})
{
    const int fin = nb <= kFinBlocks;
    launch(fin);
    if (!fin) k_sum_partials<NV, Post><<<1, kRedBlock, 0, s>>>(partial, nb, out, stop, post);
}

}  // namespace m2d
