// cgef_energy_noalloc: the area-weighted energy reduction of cgef_energy for callers that cannot
// allocate device memory. Its kernels (the per-block partial sums and the per-slice final sums) take
// their scratch from one fixed device array instead of a workspace; nothing here allocates.
//
// The array `scratch` exists once per GPU the library runs on and stays until the process ends. In
// host code its name is only a stand-in: the device address is obtained once per device with
// hipGetSymbolAddress (with that device current) and kept in `g_scratch`; every kernel argument and
// copy uses that address. Calls on one GPU share the array and leave their values in it; each call
// writes every slot it reads (the partial-sum kernel writes all blocks' slots of every slice before the
// final kernel reads them), so nothing has to start at zero. Calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <new>
#include <string>

#include "cgef_launch.hpp"

namespace cgef {

constexpr int64_t NOALLOC_PARTIALS = int64_t(1) << 20;  // per-block partial sums, all slices
constexpr int64_t NOALLOC_SLICES = 4096;                // per-slice energies
static_assert(NOALLOC_PARTIALS % 32 == 0, "the energy slots start on a 256-byte boundary");

static __device__ __align__(256) double scratch[NOALLOC_PARTIALS + NOALLOC_SLICES];

constexpr int kMaxDevices = 64;
static std::mutex g_mu;
static double *g_scratch[kMaxDevices] = {};

// The device address of `scratch` on `device` (which must be current), looked up on first use.
static double *scratch_on(int device) {
    if (device < 0 || device >= kMaxDevices) fail(CGEF_ERR_ARGUMENT, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        CGEF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) fail(CGEF_ERR_DEVICE, "hipGetSymbolAddress returned no address for the energy scratch");
        g_scratch[device] = static_cast<double *>(p);
    }
    return g_scratch[device];
}

}  // namespace cgef

using namespace cgef;

extern "C" cgef_status cgef_energy_noalloc(const cgef_grid *grid, const void *ubar, const void *vbar, const void *wbar,
                                           int64_t nbatch, int32_t dtype, double *E_out, void *stream) {
    try {
        const Grid *g = reinterpret_cast<const Grid *>(grid);
        if (!g || !ubar || !vbar || !E_out) fail(CGEF_ERR_ARGUMENT, "null grid, field or output pointer");
        if (nbatch < 1) fail(CGEF_ERR_DIMENSION, "nbatch must be >= 1");
        if (dtype != CGEF_F32 && dtype != CGEF_F64) fail(CGEF_ERR_ARGUMENT, "dtype must be CGEF_F32 or CGEF_F64");
        if (!(g->active_area > 0)) fail(CGEF_ERR_ARGUMENT, "grid has no active cells (all masked out)");
        const GeomArgs G = geom_args(*g);
        const int64_t nblk = energy_blocks(G);
        if (nbatch > NOALLOC_SLICES || nblk > NOALLOC_PARTIALS / nbatch)
            fail(CGEF_ERR_DIMENSION,
                 "cgef_energy_noalloc holds at most " + std::to_string(NOALLOC_SLICES) + " slices and " +
                     std::to_string(NOALLOC_PARTIALS) + " partial sums (slices x ceil(Nx/256)·Ny·Nz); the call needs " +
                     std::to_string(nbatch) + " x " + std::to_string(nblk));
        CGEF_HIP(hipSetDevice(g->device));
        double *base = scratch_on(g->device);
        double *partials = base, *E_dev = base + NOALLOC_PARTIALS;
        hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        launch_energy_fields(G, dtype, ubar, vbar, wbar, nbatch, partials, NOALLOC_PARTIALS, s);
        launch_energy_final(nbatch, nblk, partials, g->active_area, E_dev, s);
        CGEF_HIP(hipGetLastError());
        CGEF_HIP(hipMemcpyAsync(E_out, E_dev, (size_t)nbatch * sizeof(double), hipMemcpyDeviceToHost, s));
        CGEF_HIP(hipStreamSynchronize(s));
        return CGEF_OK;
    } catch (const Error &e) {
        set_error(e.msg);
        return e.status;
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return CGEF_ERR_ALLOC;
    } catch (const std::exception &e) {
        set_error(e.what());
        return CGEF_ERR_ARGUMENT;
    }
}
