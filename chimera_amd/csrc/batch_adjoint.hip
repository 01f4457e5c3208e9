// Batched discrete adjoint of an ensemble of independent grids of uneven sizes (the
// differentiable GRF ensemble): every forward and reverse stage is ONE persistent
// launch over a tile table spanning all members, as the batched forward ensemble does.
//
// Per member: a prescribed dt sequence (frozen), a trajectory of every stage input of
// the current segment, the fused-reverse exchange arrays and two lam buffers.  A
// segment sweep recomputes the segment (3 stage launches per step over all members),
// then per step, newest first, runs the fused reverse pair (rev_tile_a / rev_tile_b of
// adjoint_tile.hpp) for stages 3, 2, 1 over all members; each member's lam is updated in
// place.  The arithmetic per member is that of the single-grid adjoint, so a member's
// gradient is bit-identical to chimera_adj_segment_vjp's.  The recompute also keeps
// each stage's Up (the state after the flux update), which the reverse reads instead
// of recomputing the tile's fluxes.  chimera_badj_sweep runs the segments of a whole
// run from last to first in one call.
#include <algorithm>
#include <numeric>
#include <vector>

#include "adjoint_tile.hpp"
#include "common.hpp"
#include "resident.hpp"

namespace chimera {
namespace {

struct BAdjMember {
    AdjGrid G;
    size_t n;
    double *traj;               // (3 cap + 1) slots x 6 fields x n
    double *uptraj;             // 3 cap slots x 4 fields x n: Up of step s, stage st at 3s+st-1
    const double *state0[6];    // the caller's U^0 of the segment (device)
    double *lam_io[6];          // the caller's lam (device, in / out)
    double *state_out[6];       // forward: where the last step's output goes
    double *lam[3][6];          // LA, LB, lam_u0
    FusedBufs F;
    const double *dts;          // this member's dts (device)
};

struct BTile {
    int member, tx, ty;
};

// the block's share of the tile table: XCD b % 8 walks its contiguous eighth
__device__ __forceinline__ void badj_segment(int ntiles, int &begin, int &end, int &slot,
                                             int &slots) {
    const int xcd = blockIdx.x & 7;
    slot = blockIdx.x >> 3;
    slots = gridDim.x >> 3;
    begin = (int)((long long)ntiles * xcd / 8);
    end = (int)((long long)ntiles * (xcd + 1) / 8);
}

__device__ __forceinline__ double *slot_field(const BAdjMember &m, int slot, int q) {
    return m.traj + ((size_t)6 * slot + q) * m.n;
}

__device__ __forceinline__ double *slot_up(const BAdjMember &m, int slot, int q) {
    return m.uptraj + ((size_t)4 * slot + q) * m.n;
}

// stage input of step s, stage st (1..3): U^s, U^(1), U^(2)
__device__ __forceinline__ const double *stage_in(const BAdjMember &m, int s, int st, int q) {
    if (st == 1) return s == 0 ? m.state0[q] : slot_field(m, 3 * s, q);
    return slot_field(m, 3 * s + st - 1, q);
}

__device__ __forceinline__ GridView member_view(const BAdjMember &m, int s, int st) {
    GridView g;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        g.uin[q] = stage_in(m, s, st, q);
        g.u0[q] = stage_in(m, s, 1, q);
        g.uout[q] = nullptr;
    }
    g.dt = m.dts + s;
    g.mask = nullptr;
    g.nx = m.G.nx;
    g.ny = m.G.ny;
    g.px = m.G.px;
    g.py = m.G.py;
    g.dx = m.G.dx;
    g.vol = m.G.vol;
    g.gamma = m.G.gamma;
    return g;
}

// Forward stage st of step s for every member.  Segment recompute (PURE false): the
// stage output goes to the trajectory, which keeps every stage input, and Up to the Up
// trajectory.  Plain forward (PURE): the running state is the member's state_out (U^0
// read from state0), the two intermediate stage states use trajectory slots 1 and 2,
// and stage 3 overwrites state_out in place (it reads u^n only at the cell it writes).
// HLLD at three waves per SIMD (at most 168 registers a lane, as the 256-thread tiles'
// LDS allows): the compiler's 181 left two, and the two registers spilled to scratch
// cost far less than the third wave gains (forward 21% faster, the segment recompute
// inside the gradient 3%).  Rusanov and HLLE need ~135 and run three or more anyway.
template <int SOLVER, int LIMITER, int TX, int TY, bool PURE>
__global__ __launch_bounds__(STAGE_THREADS)
__attribute__((amdgpu_waves_per_eu(SOLVER == HLLD ? 3 : 1))) void badj_forward_kernel(
    const BAdjMember *__restrict__ mem, const BTile *__restrict__ tiles, int ntiles, int s, int st,
    int use_lim, int mood) {
    __shared__ double lds[TileShape<TX, TY>::LDS_DOUBLES];
    int begin, end, slot, slots;
    badj_segment(ntiles, begin, end, slot, slots);
    for (int t = begin + slot; t < end; t += slots) {
        const BTile tr = tiles[t];
        const BAdjMember &m = mem[tr.member];
        GridView g = member_view(m, s, st);
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            if (PURE) {
                const double *un = s == 0 ? m.state0[q] : m.state_out[q];
                g.uin[q] = st == 1 ? un : slot_field(m, st - 1, q);
                g.u0[q] = un;
                g.uout[q] = st == 3 ? m.state_out[q] : slot_field(m, st, q);
            } else {
                g.uout[q] = slot_field(m, 3 * s + st, q);
            }
        }
        if (PURE) {
            stage_tile<SOLVER, LIMITER, TX, TY>(g, st, use_lim != 0, mood != 0, tr.tx, tr.ty, lds);
        } else {
            StageStore store{};
#pragma unroll
            for (int q = 0; q < 4; ++q) store.Up[q] = slot_up(m, 3 * s + st - 1, q);
            stage_tile<SOLVER, LIMITER, TX, TY, STORE_UP>(g, st, use_lim != 0, mood != 0, tr.tx,
                                                          tr.ty, lds, &store);
        }
        __syncthreads();   // the next tile reuses LDS
    }
}

// lam buffers of stage st: lout -> lin (stage 3: lam_io -> LB, 2: LB -> LA, 1: LA -> lam_io)
__device__ __forceinline__ void stage_lams(const BAdjMember &m, int st, const double *lout[6],
                                           double *lin[6]) {
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        lout[q] = st == 3 ? m.lam_io[q] : (st == 2 ? m.lam[1][q] : m.lam[0][q]);
        lin[q] = st == 3 ? m.lam[1][q] : (st == 2 ? m.lam[0][q] : m.lam_io[q]);
    }
}

// Two waves per SIMD (at most 256 registers a lane): the HLLD face reverse needs ~300,
// and held at one wave per SIMD it left the VALU idle 61% of the launch; spilling the
// excess to scratch at two waves is 8% faster over the whole HLLD gradient.  The other
// solvers fit in 256 without spilling.
template <int SOLVER, int LIMITER, int TX, int TY>
__global__ __launch_bounds__(TX *TY) __attribute__((amdgpu_waves_per_eu(2, 2))) void badj_rev_a_kernel(const BAdjMember *__restrict__ mem,
                                                           const BTile *__restrict__ tiles,
                                                           int ntiles, int s, int st) {
    __shared__ double lds[RevShapeA<TX, TY>::LDS_DOUBLES];
    int begin, end, slot, slots;
    badj_segment(ntiles, begin, end, slot, slots);
    for (int t = begin + slot; t < end; t += slots) {
        const BTile tr = tiles[t];
        const BAdjMember &m = mem[tr.member];
        const GridView g = member_view(m, s, st);
        const double *lout[6], *up[4];
        double *lin[6], *lu0[6];
        stage_lams(m, st, lout, lin);
#pragma unroll
        for (int q = 0; q < 6; ++q) lu0[q] = m.lam[2][q];
#pragma unroll
        for (int q = 0; q < 4; ++q) up[q] = slot_up(m, 3 * s + st - 1, q);
        double dt = m.dts[s];
        rev_tile_a<SOLVER, LIMITER, TX, TY, true>(g, m.G, m.F, up, lout, lin, lu0, st, dt, tr.tx,
                                                  tr.ty, lds);
        __syncthreads();
    }
}

// Four waves per SIMD (at most 128 registers): with one field round of loads in flight
// (rev_tile_b without EARLY_TAIL) the kernel needs 131 and spills two; the fourth wave
// hides the loads' latency better than the prefetch did (the HLLD gradient 3% faster).
template <int LIMITER, int TX, int TY>
__global__ __launch_bounds__(TX *TY) __attribute__((amdgpu_waves_per_eu(4))) void badj_rev_b_kernel(const BAdjMember *__restrict__ mem,
                                                           const BTile *__restrict__ tiles,
                                                           int ntiles, int s, int st) {
    __shared__ double lds[RevShapeB<TX, TY>::LDS_DOUBLES];
    int begin, end, slot, slots;
    badj_segment(ntiles, begin, end, slot, slots);
    for (int t = begin + slot; t < end; t += slots) {
        const BTile tr = tiles[t];
        const BAdjMember &m = mem[tr.member];
        const double *U[6], *lout[6], *lu0[6];
        double *lin[6];
        stage_lams(m, st, lout, lin);
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            U[q] = stage_in(m, s, st, q);
            lu0[q] = m.lam[2][q];
        }
        rev_tile_b<LIMITER, TX, TY, false>(m.G, U, m.F, lin, st == 1 ? lu0 : nullptr, tr.tx,
                                               tr.ty, lds);
        __syncthreads();
    }
}

// tile shapes of every pass (forward, reverse A, reverse B): 0 = 16 x 16, 1 = 32 x 8
constexpr int BADJ_NUM_SHAPES = 2;
constexpr int BADJ_SHAPES[BADJ_NUM_SHAPES][2] = {{16, 16}, {32, 8}};

using BFwdFn = void (*)(const BAdjMember *, const BTile *, int, int, int, int, int);
using BRevAFn = void (*)(const BAdjMember *, const BTile *, int, int, int);
using BRevBFn = void (*)(const BAdjMember *, const BTile *, int, int, int);

template <int S, int TX, int TY, bool PURE>
constexpr BFwdFn fwd_of(int limiter) {
    return limiter == LIM_MC ? badj_forward_kernel<S, LIM_MC, TX, TY, PURE>
                             : badj_forward_kernel<S, LIM_VAN_LEER, TX, TY, PURE>;
}
template <int TX, int TY, bool PURE>
BFwdFn fwd_shape(int solver, int limiter) {
    return solver == RUSANOV ? fwd_of<RUSANOV, TX, TY, PURE>(limiter)
         : solver == HLLE    ? fwd_of<HLLE, TX, TY, PURE>(limiter)
                             : fwd_of<HLLD, TX, TY, PURE>(limiter);
}
template <bool PURE>
BFwdFn fwd_fn(int solver, int limiter, int shape) {
    return shape == 0 ? fwd_shape<16, 16, PURE>(solver, limiter)
                      : fwd_shape<32, 8, PURE>(solver, limiter);
}

template <int S, int TX, int TY>
constexpr BRevAFn reva_of(int limiter) {
    return limiter == LIM_MC ? badj_rev_a_kernel<S, LIM_MC, TX, TY>
                             : badj_rev_a_kernel<S, LIM_VAN_LEER, TX, TY>;
}
BRevAFn reva_fn(int solver, int limiter, int shape) {
    if (shape == 0)
        return solver == RUSANOV ? reva_of<RUSANOV, 16, 16>(limiter)
             : solver == HLLE    ? reva_of<HLLE, 16, 16>(limiter)
                                 : reva_of<HLLD, 16, 16>(limiter);
    return solver == RUSANOV ? reva_of<RUSANOV, 32, 8>(limiter)
         : solver == HLLE    ? reva_of<HLLE, 32, 8>(limiter)
                             : reva_of<HLLD, 32, 8>(limiter);
}
BRevBFn revb_fn(int limiter, int shape) {
    if (shape == 0)
        return limiter == LIM_MC ? badj_rev_b_kernel<LIM_MC, 16, 16>
                                 : badj_rev_b_kernel<LIM_VAN_LEER, 16, 16>;
    return limiter == LIM_MC ? badj_rev_b_kernel<LIM_MC, 32, 8>
                             : badj_rev_b_kernel<LIM_VAN_LEER, 32, 8>;
}

}  // namespace
}  // namespace chimera

using namespace chimera;

struct chimera_badj {
    int members = 0;
    chimera_scheme_t sch{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::vector<BAdjMember> hmem;
    BAdjMember *dmem = nullptr;
    double *arena = nullptr;       // per member: lam buffers and exchange arrays
    std::vector<double *> traj;    // per member trajectory (grown on demand)
    std::vector<double *> uptraj;  // per member Up trajectory
    int traj_cap = 0;              // steps the trajectories hold
    double *dts = nullptr;         // members x dt_cap
    int dt_cap = 0;
    BTile *dtiles[BADJ_NUM_SHAPES] = {};   // tile tables of the shapes
    int ntiles[BADJ_NUM_SHAPES] = {};
    int shape_f = 0, shape_a = 1, shape_b = 1;   // forward, reverse A, reverse B
    int cu_count = 0;
    long long cells = 0;
    // the whole-run sweep: one member table per segment
    BAdjMember *dseg = nullptr;
    int dseg_cap = 0;              // segments the tables hold
    // the last segment of a run recorded by chimera_badj_record for the next sweep
    struct Recorded {
        bool valid = false;
        std::vector<const double *> state0;   // members x 6
        std::vector<double> dts;              // members x nsteps
        int nsteps = 0, K = 0, nseg = 0;
    } rec;
};

namespace {

// Blocks of a persistent pass over a table of ntiles tiles: 16 x what the CUs hold at
// once (capped at the tile count), as the batched forward ensemble sizes its launch --
// blocks that start whenever a slot frees up keep the blocks sharing a CU out of phase.
int badj_grid(chimera_badj *b, const void *fn, int threads, int ntiles, int &grid) {
    int occ = 0;
    CH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, threads, 0));
    long long g = (long long)b->cu_count * std::max(occ, 1) * 16;
    g = std::min(g, ((long long)ntiles + 7) / 8 * 8);
    grid = (int)std::max(8ll, g / 8 * 8);
    return CHIMERA_OK;
}

int badj_tables(chimera_badj *b) {
    std::vector<int> order(b->members);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return b->hmem[x].n > b->hmem[y].n;
    });
    for (int sh = 0; sh < BADJ_NUM_SHAPES; ++sh) {
        std::vector<BTile> t;
        for (int m : order) {
            const int tx = (b->hmem[m].G.nx + BADJ_SHAPES[sh][0] - 1) / BADJ_SHAPES[sh][0];
            const int ty = (b->hmem[m].G.ny + BADJ_SHAPES[sh][1] - 1) / BADJ_SHAPES[sh][1];
            for (int y = 0; y < ty; ++y)
                for (int x = 0; x < tx; ++x) t.push_back({m, x, y});
        }
        b->ntiles[sh] = (int)t.size();
        CH_HIP(hipMalloc(&b->dtiles[sh], t.size() * sizeof(BTile)));
        CH_HIP(hipMemcpy(b->dtiles[sh], t.data(), t.size() * sizeof(BTile), hipMemcpyHostToDevice));
    }
    return CHIMERA_OK;
}

// per-call pointers (state0, lam, state_out, dts) into the device member table
int badj_upload(chimera_badj *b, const double *const *state0, double *const *lam,
                double *const *state_out) {
    for (int m = 0; m < b->members; ++m) {
        BAdjMember &h = b->hmem[m];
        for (int q = 0; q < 6; ++q) {
            h.state0[q] = state0 ? state0[6 * m + q] : nullptr;
            h.lam_io[q] = lam ? lam[6 * m + q] : nullptr;
            h.state_out[q] = state_out ? state_out[6 * m + q] : nullptr;
        }
        h.traj = b->traj[m];
        h.uptraj = b->uptraj[m];
        h.dts = b->dts + (size_t)m * b->dt_cap;
    }
    CH_HIP(hipStreamSynchronize(b->stream));
    CH_HIP(hipMemcpy(b->dmem, b->hmem.data(), b->hmem.size() * sizeof(BAdjMember),
                     hipMemcpyHostToDevice));
    return CHIMERA_OK;
}

// dts (host, members x nsteps) to the device, and trajectories for traj_steps steps
int badj_prepare(chimera_badj *b, const double *dt_in, int nsteps, int traj_steps) {
    CH_HIP(hipStreamSynchronize(b->stream));
    if (nsteps > b->dt_cap) {
        (void)hipFree(b->dts);
        b->dts = nullptr;
        b->dt_cap = 0;
        CH_HIP(hipMalloc(&b->dts, (size_t)b->members * nsteps * sizeof(double)));
        b->dt_cap = nsteps;
    }
    std::vector<double> h((size_t)b->members * b->dt_cap, 0.0);
    for (int m = 0; m < b->members; ++m)
        for (int s = 0; s < nsteps; ++s) h[(size_t)m * b->dt_cap + s] = dt_in[(size_t)m * nsteps + s];
    CH_HIP(hipMemcpy(b->dts, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    if (traj_steps > b->traj_cap) {
        b->rec.valid = false;   // a recorded segment does not survive a reallocation
        for (auto *v : {&b->traj, &b->uptraj})
            for (auto &p : *v) {
                (void)hipFree(p);
                p = nullptr;
            }
        b->traj_cap = 0;
        for (int m = 0; m < b->members; ++m) {
            const size_t n = b->hmem[m].n;
            const size_t bytes = (size_t)(3 * traj_steps + 1) * 6 * n * sizeof(double);
            const size_t ubytes = (size_t)3 * traj_steps * 4 * n * sizeof(double);
            if (hipMalloc(&b->traj[m], bytes) != hipSuccess ||
                hipMalloc(&b->uptraj[m], ubytes) != hipSuccess) {
                set_error("batched adjoint: trajectory of member %d (%zu bytes) failed", m,
                          bytes + ubytes);
                return CHIMERA_ENOMEM;
            }
        }
        b->traj_cap = traj_steps;
    }
    return CHIMERA_OK;
}

// nsteps forward steps of every member of the table mem from its state0.  pure: into
// state_out, keeping nothing; else the segment recompute, keeping every stage input
// and every stage's Up in the trajectories.
int badj_forward(chimera_badj *b, const BAdjMember *mem, int nsteps, bool pure) {
    const BFwdFn fn = pure ? fwd_fn<true>(b->sch.riemann_solver, b->sch.slope_limiter, b->shape_f)
                           : fwd_fn<false>(b->sch.riemann_solver, b->sch.slope_limiter, b->shape_f);
    int grid;
    CH_TRY(badj_grid(b, (const void *)fn, STAGE_THREADS, b->ntiles[b->shape_f], grid));
    for (int s = 0; s < nsteps; ++s)
        for (int stage = 1; stage <= 3; ++stage) {
            hipLaunchKernelGGL(fn, dim3(grid), dim3(STAGE_THREADS), 0, b->stream, mem,
                               b->dtiles[b->shape_f], b->ntiles[b->shape_f], s, stage,
                               b->sch.use_slope_limiting, b->sch.fallback_2_mood);
            CH_TRY(launch_check("badj_forward_kernel"));
        }
    return CHIMERA_OK;
}

// The reverse of a recomputed segment of nsteps steps for every member of the table mem,
// on the main stream: per step, newest first, the fused pair for stages 3, 2, 1.
int badj_reverse(chimera_badj *b, const BAdjMember *mem, int nsteps) {
    const BRevAFn fa = reva_fn(b->sch.riemann_solver, b->sch.slope_limiter, b->shape_a);
    const BRevBFn fb = revb_fn(b->sch.slope_limiter, b->shape_b);
    const int ta = BADJ_SHAPES[b->shape_a][0] * BADJ_SHAPES[b->shape_a][1];
    const int tb = BADJ_SHAPES[b->shape_b][0] * BADJ_SHAPES[b->shape_b][1];
    int ga, gb;
    CH_TRY(badj_grid(b, (const void *)fa, ta, b->ntiles[b->shape_a], ga));
    CH_TRY(badj_grid(b, (const void *)fb, tb, b->ntiles[b->shape_b], gb));
    for (int s = nsteps - 1; s >= 0; --s)
        for (int st = 3; st >= 1; --st) {
            hipLaunchKernelGGL(fa, dim3(ga), dim3(ta), 0, b->stream, mem, b->dtiles[b->shape_a],
                               b->ntiles[b->shape_a], s, st);
            hipLaunchKernelGGL(fb, dim3(gb), dim3(tb), 0, b->stream, mem, b->dtiles[b->shape_b],
                               b->ntiles[b->shape_b], s, st);
            CH_TRY(launch_check("batched adjoint reverse stage"));
        }
    return CHIMERA_OK;
}

int ensure_sweep(chimera_badj *b, int nseg) {
    if (nseg > b->dseg_cap) {
        (void)hipFree(b->dseg);
        b->dseg = nullptr;
        b->dseg_cap = 0;
        CH_HIP(hipMalloc(&b->dseg, (size_t)nseg * b->members * sizeof(BAdjMember)));
        b->dseg_cap = nseg;
    }
    return CHIMERA_OK;
}

}  // namespace

extern "C" int chimera_badj_destroy(chimera_badj_t *b) {
    if (!b) return CHIMERA_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (auto p : b->traj) (void)hipFree(p);
    for (auto p : b->uptraj) (void)hipFree(p);
    for (auto p : b->dtiles) (void)hipFree(p);
    (void)hipFree(b->arena);
    (void)hipFree(b->dmem);
    (void)hipFree(b->dseg);
    (void)hipFree(b->dts);
    if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return CHIMERA_OK;
}

extern "C" int chimera_badj_create(chimera_badj_t **out, int members, const int *nx, const int *ny,
                                   const double *dx, const double *gamma, const chimera_bc_t *bc,
                                   const chimera_scheme_t *scheme, void *stream) {
    CH_REQUIRE(out && nx && ny && dx && gamma && bc && scheme, "NULL argument");
    CH_REQUIRE(members >= 1, "members must be >= 1");
    CH_TRY(check_scheme(scheme));
    chimera_badj *b = new chimera_badj();
    b->members = members;
    b->sch = *scheme;
    b->hmem.resize(members);
    b->traj.assign(members, nullptr);
    b->uptraj.assign(members, nullptr);
    size_t total = 0;
    for (int m = 0; m < members; ++m) {
        if (!(nx[m] >= 4 && ny[m] >= 4 && dx[m] > 0.0 && gamma[m] > 1.0 &&
              (long long)nx[m] * ny[m] <= MAX_GRID_CELLS)) {
            set_error("member %d: need nx, ny >= 4, nx * ny <= %lld, dx > 0 and gamma > 1", m,
                      MAX_GRID_CELLS);
            delete b;
            return CHIMERA_EINVAL;
        }
        int px = 0, py = 0;
        const int rc = check_bc(&bc[m], &px, &py);
        if (rc != CHIMERA_OK) {
            delete b;
            return rc;
        }
        BAdjMember &h = b->hmem[m];
        h = BAdjMember{};
        h.G = AdjGrid{nx[m], ny[m], px, py, dx[m], dx[m] * dx[m], gamma[m],
                      scheme->use_slope_limiting, scheme->fallback_2_mood};
        h.n = (size_t)nx[m] * ny[m];
        total += h.n * (18 + FusedBufs::ARRAYS);
        b->cells += (long long)h.n;
    }
    if (stream) {
        b->stream = (hipStream_t)stream;
    } else {
        if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
            delete b;
            set_error("hipStreamCreate failed");
            return CHIMERA_EHIP;
        }
        b->own_stream = true;
    }
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess ||
        hipMalloc(&b->arena, total * sizeof(double)) != hipSuccess ||
        hipMalloc(&b->dmem, members * sizeof(BAdjMember)) != hipSuccess) {
        set_error("batched adjoint: device set-up (%zu doubles of workspace) failed", total);
        chimera_badj_destroy(b);
        return CHIMERA_ENOMEM;
    }
    b->cu_count = prop.multiProcessorCount;
    double *p = b->arena;
    for (auto &h : b->hmem) {
        for (auto &row : h.lam)
            for (auto &x : row) {
                x = p;
                p += h.n;
            }
        h.F = FusedBufs{p, h.n};
        p += (size_t)FusedBufs::ARRAYS * h.n;
    }
    int rc = badj_tables(b);
    if (rc != CHIMERA_OK) {
        chimera_badj_destroy(b);
        return rc;
    }
    *out = b;
    return CHIMERA_OK;
}

// Every member nsteps steps forward with its prescribed dts (dt_in: members x nsteps,
// host), from state_in to state_out (members x 6 device pointers; may be equal).
extern "C" int chimera_badj_forward(chimera_badj_t *b, const double *const *state_in,
                                    const double *dt_in, int nsteps, double *const *state_out) {
    CH_REQUIRE(b && state_in && dt_in && state_out, "NULL argument");
    CH_REQUIRE(nsteps >= 1, "nsteps must be >= 1");
    for (long long k = 0; k < (long long)b->members * nsteps; ++k)
        CH_REQUIRE(dt_in[k] > 0.0, "dt_in[%lld] = %g is not positive", k, dt_in[k]);
    b->rec.valid = false;   // slots 1 and 2 of the trajectory are overwritten
    CH_TRY(badj_prepare(b, dt_in, nsteps, 1));   // slots 1 and 2 of one step
    CH_TRY(badj_upload(b, state_in, nullptr, state_out));
    CH_TRY(badj_forward(b, b->dmem, nsteps, true));
    CH_HIP(hipStreamSynchronize(b->stream));
    return CHIMERA_OK;
}

// Reverse sweep over one checkpoint segment for every member: recompute the segment
// from state0 (members x 6 device pointers) with the dts dt_in (members x nsteps,
// host), then lam <- (dU^{nsteps}/dU^0)^T lam in place (members x 6 device pointers).
extern "C" int chimera_badj_segment_vjp(chimera_badj_t *b, const double *const *state0,
                                        const double *dt_in, int nsteps, double *const *lam) {
    CH_REQUIRE(b && state0 && dt_in && lam, "NULL argument");
    CH_REQUIRE(nsteps >= 1, "nsteps must be >= 1");
    for (long long k = 0; k < (long long)b->members * nsteps; ++k)
        CH_REQUIRE(dt_in[k] > 0.0, "dt_in[%lld] = %g is not positive", k, dt_in[k]);
    b->rec.valid = false;
    CH_TRY(badj_prepare(b, dt_in, nsteps, nsteps));
    CH_TRY(badj_upload(b, state0, lam, nullptr));
    CH_TRY(badj_forward(b, b->dmem, nsteps, false));
    CH_TRY(badj_reverse(b, b->dmem, nsteps));
    CH_HIP(hipStreamSynchronize(b->stream));
    return CHIMERA_OK;
}

// Forward of the last checkpoint segment of a run for every member (nsteps <= seg_steps
// steps from state0, members x 6 device pointers, dts dt_in members x nsteps, host),
// keeping the trajectories chimera_badj_sweep of a run of nseg segments would
// recompute for that segment; the final states go to state_out (members x 6).
extern "C" int chimera_badj_record(chimera_badj_t *b, const double *const *state0,
                                   const double *dt_in, int nsteps, int seg_steps, int nseg,
                                   double *const *state_out) {
    CH_REQUIRE(b && state0 && dt_in && state_out, "NULL argument");
    CH_REQUIRE(nsteps >= 1 && nseg >= 1 && seg_steps >= nsteps,
               "need 1 <= nsteps <= seg_steps and nseg >= 1 (got %d, %d, %d)", nsteps, seg_steps,
               nseg);
    for (long long k = 0; k < (long long)b->members * nsteps; ++k)
        CH_REQUIRE(dt_in[k] > 0.0, "dt_in[%lld] = %g is not positive", k, dt_in[k]);
    b->rec.valid = false;
    const int K = nseg > 1 ? seg_steps : nsteps;
    CH_TRY(badj_prepare(b, dt_in, nsteps, K));
    CH_TRY(badj_upload(b, state0, nullptr, nullptr));
    CH_TRY(badj_forward(b, b->dmem, nsteps, false));
    for (int m = 0; m < b->members; ++m) {
        const BAdjMember &h = b->hmem[m];
        for (int q = 0; q < 6; ++q)
            CH_HIP(hipMemcpyAsync(state_out[6 * (size_t)m + q],
                                  b->traj[m] + ((size_t)6 * (3 * nsteps) + q) * h.n,
                                  h.n * sizeof(double), hipMemcpyDeviceToDevice, b->stream));
    }
    CH_HIP(hipStreamSynchronize(b->stream));
    b->rec.valid = true;
    b->rec.state0.assign(state0, state0 + 6 * (size_t)b->members);
    b->rec.dts.assign(dt_in, dt_in + (size_t)b->members * nsteps);
    b->rec.nsteps = nsteps;
    b->rec.K = K;
    b->rec.nseg = nseg;
    return CHIMERA_OK;
}

// The whole reverse sweep of an nsteps run checkpointed every seg_steps steps for every
// member: states holds nseg x members x 6 device pointers (checkpoint k of member m at
// 6 (k members + m)), dt_in members x nsteps (host), lam members x 6 device pointers
// (in place).  The segments are recomputed and reversed from last to first in one
// call, with one member table per segment.  (Overlapping a recompute with the
// previous segment's reverse on a second stream gains nothing here: an ensemble's
// reverse passes already hold every CU, and two segments' trajectories of a
// 64-member ensemble would not fit in HBM.)
extern "C" int chimera_badj_sweep(chimera_badj_t *b, const double *const *states,
                                  const double *dt_in, int nsteps, int seg_steps,
                                  double *const *lam) {
    CH_REQUIRE(b && states && dt_in && lam, "NULL argument");
    CH_REQUIRE(nsteps >= 1 && seg_steps >= 1, "nsteps and seg_steps must be >= 1");
    for (long long k = 0; k < (long long)b->members * nsteps; ++k)
        CH_REQUIRE(dt_in[k] > 0.0, "dt_in[%lld] = %g is not positive", k, dt_in[k]);
    const int nseg = (nsteps + seg_steps - 1) / seg_steps;
    for (long long k = 0; k < 6LL * nseg * b->members; ++k)
        CH_REQUIRE(states[k], "states[%lld] is NULL", k);
    const int K = std::min(seg_steps, nsteps);
    CH_TRY(badj_prepare(b, dt_in, nsteps, K));
    CH_TRY(ensure_sweep(b, nseg));
    // the last segment recorded by chimera_badj_record for exactly this run (checked
    // after the allocations, which drop a recording when they reallocate)?
    const int last_len = nsteps - (nseg - 1) * seg_steps;
    bool have_last = b->rec.valid && b->rec.nseg == nseg && b->rec.K == K &&
                     b->rec.nsteps == last_len;
    for (int m = 0; have_last && m < b->members; ++m) {
        for (int q = 0; q < 6; ++q)
            have_last = have_last && b->rec.state0[6 * (size_t)m + q] ==
                                         states[6 * ((size_t)(nseg - 1) * b->members + m) + q];
        for (int s = 0; s < last_len; ++s)
            have_last = have_last && b->rec.dts[(size_t)m * last_len + s] ==
                                         dt_in[(size_t)m * nsteps + (size_t)(nseg - 1) * seg_steps + s];
    }
    b->rec.valid = false;
    // one member table per segment: its checkpoint and dt offset
    std::vector<BAdjMember> tabs((size_t)nseg * b->members);
    for (int k = 0; k < nseg; ++k)
        for (int m = 0; m < b->members; ++m) {
            BAdjMember h = b->hmem[m];
            for (int q = 0; q < 6; ++q) {
                h.state0[q] = states[6 * ((size_t)k * b->members + m) + q];
                h.lam_io[q] = lam[6 * (size_t)m + q];
                h.state_out[q] = nullptr;
            }
            h.traj = b->traj[m];
            h.uptraj = b->uptraj[m];
            h.dts = b->dts + (size_t)m * b->dt_cap + (size_t)k * seg_steps;
            tabs[(size_t)k * b->members + m] = h;
        }
    CH_HIP(hipStreamSynchronize(b->stream));
    CH_HIP(hipMemcpy(b->dseg, tabs.data(), tabs.size() * sizeof(BAdjMember), hipMemcpyHostToDevice));
    for (int k = nseg - 1; k >= 0; --k) {
        const BAdjMember *tab = b->dseg + (size_t)k * b->members;
        const int len = std::min(seg_steps, nsteps - k * seg_steps);
        if (!(have_last && k == nseg - 1)) CH_TRY(badj_forward(b, tab, len, false));
        CH_TRY(badj_reverse(b, tab, len));
    }
    CH_HIP(hipStreamSynchronize(b->stream));
    return CHIMERA_OK;
}

// Tile shapes of the two reverse passes and of the forward stages: 0 = 16 x 16,
// 1 = 32 x 8 (-1 keeps).
extern "C" int chimera_badj_set_tiles(chimera_badj_t *b, int shape_a, int shape_b, int shape_f) {
    CH_REQUIRE(b, "NULL batched adjoint");
    for (int v : {shape_a, shape_b, shape_f})
        CH_REQUIRE(v >= -1 && v < BADJ_NUM_SHAPES, "tile shapes must be -1, 0 or 1, got %d", v);
    if (shape_a >= 0) b->shape_a = shape_a;
    if (shape_b >= 0) b->shape_b = shape_b;
    if (shape_f >= 0) b->shape_f = shape_f;
    return CHIMERA_OK;
}

// Tile autotuner: times one forward step of every member from state0 (members x 6
// device pointers, dts dt_in: members x 1, host) per forward shape, keeps the fastest,
// recomputes that step, then times one reverse step per (A, B) shape pair and keeps the
// fastest.  reps timed repetitions each after one untimed one.  Writes scratch arrays
// of its own (the caller's are only read); best (may be NULL) gets the chosen
// {forward, A, B} shapes, ms (may be NULL) the mean ms of the forward shapes 0, 1 and
// of the pairs (A, B) = (0,0), (0,1), (1,0), (1,1).
extern "C" int chimera_badj_autotune(chimera_badj_t *b, const double *const *state0,
                                     const double *dt_in, int reps, int *best, double *ms) {
    CH_REQUIRE(b && state0 && dt_in, "NULL argument");
    CH_REQUIRE(reps >= 1, "reps must be >= 1");
    for (int m = 0; m < b->members; ++m)
        CH_REQUIRE(dt_in[m] > 0.0, "dt_in[%d] = %g is not positive", m, dt_in[m]);
    b->rec.valid = false;
    CH_TRY(badj_prepare(b, dt_in, 1, 1));
    // scratch: 6 arrays per member, the pure forward's output and then lam
    size_t total = 0;
    for (const auto &h : b->hmem) total += 6 * h.n;
    double *scratch = nullptr;
    CH_HIP(hipMalloc(&scratch, total * sizeof(double)));
    std::vector<double *> sp((size_t)6 * b->members);
    {
        double *p = scratch;
        for (int m = 0; m < b->members; ++m)
            for (int q = 0; q < 6; ++q) {
                sp[6 * (size_t)m + q] = p;
                p += b->hmem[m].n;
            }
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = CHIMERA_OK;
    auto timed = [&](auto &&body, double &out) -> int {
        CH_TRY(body());   // untimed: code load, caches
        CH_HIP(hipEventRecord(e0, b->stream));
        for (int r = 0; r < reps; ++r) CH_TRY(body());
        CH_HIP(hipEventRecord(e1, b->stream));
        CH_HIP(hipEventSynchronize(e1));
        float t = 0.0f;
        CH_HIP(hipEventElapsedTime(&t, e0, e1));
        out = (double)t / reps;
        return CHIMERA_OK;
    };
    auto run = [&]() -> int {
        CH_HIP(hipEventCreate(&e0));
        CH_HIP(hipEventCreate(&e1));
        CH_HIP(hipMemsetAsync(scratch, 0, total * sizeof(double), b->stream));
        double t[2 + 2 * BADJ_NUM_SHAPES] = {};
        CH_TRY(badj_upload(b, state0, nullptr, sp.data()));
        for (int f = 0; f < BADJ_NUM_SHAPES; ++f) {
            b->shape_f = f;
            CH_TRY(timed([&] { return badj_forward(b, b->dmem, 1, true); }, t[f]));
        }
        b->shape_f = t[1] < t[0] ? 1 : 0;
        CH_TRY(badj_upload(b, state0, sp.data(), nullptr));
        CH_TRY(badj_forward(b, b->dmem, 1, false));
        int ba = b->shape_a, bb = b->shape_b;
        double tbest = 0.0;
        for (int a = 0; a < BADJ_NUM_SHAPES; ++a)
            for (int c = 0; c < BADJ_NUM_SHAPES; ++c) {
                b->shape_a = a;
                b->shape_b = c;
                double &tt = t[2 + 2 * a + c];
                CH_TRY(timed([&] { return badj_reverse(b, b->dmem, 1); }, tt));
                if ((a == 0 && c == 0) || tt < tbest) {
                    tbest = tt;
                    ba = a;
                    bb = c;
                }
            }
        b->shape_a = ba;
        b->shape_b = bb;
        if (best) {
            best[0] = b->shape_f;
            best[1] = ba;
            best[2] = bb;
        }
        if (ms)
            for (int k = 0; k < 2 + 2 * BADJ_NUM_SHAPES; ++k) ms[k] = t[k];
        return CHIMERA_OK;
    };
    rc = run();
    (void)hipStreamSynchronize(b->stream);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(scratch);
    return rc;
}

extern "C" void *chimera_badj_stream(chimera_badj_t *b) { return b ? (void *)b->stream : nullptr; }
