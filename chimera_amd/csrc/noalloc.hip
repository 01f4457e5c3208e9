// One SSP-RK3 step of a caller-owned grid without any device allocation
// (chimera_step_noalloc): the two intermediate stage states and the device dt live in
// a library-owned static device array, so the entry point suits callers that cannot
// allocate (code under HIP graph capture, tight serving loops).  The kernels are the
// fused stage kernels of the resident driver, so the result has the bits of
// chimera_adj_forward_step and of chimera_ctx_steps with the same dt.
#include <mutex>

#include "common.hpp"
#include "resident.hpp"

namespace chimera {

// Largest grid the scratch holds, and its layout: stage states U1 and U2 (6 fields of
// NOALLOC_MAX_CELLS doubles each), then the step's dt.
constexpr long long NOALLOC_MAX_CELLS = 512ll * 512ll;
constexpr size_t NOALLOC_DT = 12 * (size_t)NOALLOC_MAX_CELLS;
constexpr size_t NOALLOC_DOUBLES = NOALLOC_DT + 32;   // dt slot padded to 256 bytes

// Each GPU holds its own copy until the process ends.  Host code never uses this name
// as an address: it uses the per-device pointer from hipGetSymbolAddress below.
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

__global__ void noalloc_set_dt_kernel(double *dt_slot, double dt) { *dt_slot = dt; }

namespace {

constexpr int MAX_DEVICES = 64;
struct DeviceScratch {
    double *ptr = nullptr;   // device address of `scratch` on this device
    int cus = 0;
};
std::mutex g_mutex;
DeviceScratch g_dev[MAX_DEVICES];

// The current device's scratch address (looked up once per device) and CU count.
int device_scratch(DeviceScratch *out) {
    int dev = 0;
    CH_HIP(hipGetDevice(&dev));
    CH_REQUIRE(dev >= 0 && dev < MAX_DEVICES, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_mutex);
    DeviceScratch &d = g_dev[dev];
    if (!d.ptr) {
        void *p = nullptr;
        CH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        CH_REQUIRE(p != nullptr, "no device address for the step scratch");
        int cus = 0;
        CH_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        d.cus = cus;
        d.ptr = static_cast<double *>(p);
    }
    *out = d;
    return CHIMERA_OK;
}

}  // namespace
}  // namespace chimera

using namespace chimera;

extern "C" int chimera_step_noalloc(const double *const *state_in, double dt,
                                    double *const *state_out, int nx, int ny, double dx,
                                    double gamma, const chimera_bc_t *bc,
                                    const chimera_scheme_t *scheme, void *stream) {
    CH_REQUIRE(state_in && state_out, "NULL argument");
    for (int q = 0; q < 6; ++q)
        CH_REQUIRE(state_in[q] && state_out[q], "state pointer %d is NULL", q);
    CH_REQUIRE(nx >= 4 && ny >= 4, "grid must be at least 4 x 4 (got %d x %d)", nx, ny);
    CH_REQUIRE((long long)nx * ny <= NOALLOC_MAX_CELLS,
               "grid of %d x %d cells exceeds the %lld cells of the step scratch", nx, ny,
               NOALLOC_MAX_CELLS);
    CH_REQUIRE(dx > 0.0 && gamma > 1.0, "dx must be > 0 and gamma > 1");
    CH_REQUIRE(dt > 0.0, "dt must be > 0");
    int px = 0, py = 0;
    CH_TRY(check_bc(bc, &px, &py));
    CH_TRY(check_scheme(scheme));
    DeviceScratch ds;
    CH_TRY(device_scratch(&ds));
    hipStream_t s = static_cast<hipStream_t>(stream);

    const size_t n = (size_t)nx * ny;
    double *U1[6], *U2[6];
    for (int q = 0; q < 6; ++q) {
        U1[q] = ds.ptr + (size_t)q * n;
        U2[q] = ds.ptr + (size_t)(6 + q) * n;
    }
    double *dt_slot = ds.ptr + NOALLOC_DT;
    hipLaunchKernelGGL(noalloc_set_dt_kernel, dim3(1), dim3(1), 0, s, dt_slot, dt);
    CH_TRY(launch_check("noalloc_set_dt_kernel"));

    // tile shape as a new resident context picks it
    const long long tiles16 = (long long)((nx + 31) / 32) * ((ny + 15) / 16);
    const int shape = (scheme->riemann_solver != HLLD && ds.cus > 0 && tiles16 >= ds.cus) ? 4 : 0;
    const TileDims td = RES_TILE_SHAPES[shape];
    StageLaunch L{};
    L.g.dt = dt_slot;
    L.g.mask = nullptr;
    L.g.nx = nx;
    L.g.ny = ny;
    L.g.px = px;
    L.g.py = py;
    L.g.dx = dx;
    L.g.vol = dx * dx;
    L.g.gamma = gamma;
    L.use_lim = scheme->use_slope_limiting;
    L.mood = scheme->fallback_2_mood;
    L.solver = scheme->riemann_solver;
    L.limiter = scheme->slope_limiter;
    L.shape = shape;
    L.tiles_x = (nx + td.tx - 1) / td.tx;
    L.ntiles = L.tiles_x * ((ny + td.ty - 1) / td.ty);
    L.grid = (L.ntiles + 7) / 8 * 8;
    // stage k reads io[k][0] (with its halo) and io[k][1] (u^n, own cell only) and
    // writes io[k][2]; stage 3 reads u^n and writes the output cell by cell in the same
    // thread, so state_out may be state_in
    const double *const *in = state_in;
    for (int k = 0; k < 3; ++k) {
        for (int q = 0; q < 6; ++q) {
            L.g.uin[q] = k == 0 ? in[q] : (k == 1 ? U1[q] : U2[q]);
            L.g.u0[q] = in[q];
            L.g.uout[q] = k == 0 ? U1[q] : (k == 1 ? U2[q] : state_out[q]);
        }
        L.stage = k + 1;
        CH_TRY(launch_stage(L, s));
    }
    return CHIMERA_OK;
}
