// gfx950 kernels that measure what keeping the blocked red-black solver
// resident between its K-sweep steps would cost, against relaunching it.
//
// A step is the blocked kernel's shape on the 1026 x 1026 grid: a 16 x 16
// grid of workgroups of 768 threads, each writing its 64 x 64 FP64 output
// tile (32 KB).  A resident solver would replace the kernel boundary by one
// of two waits, both built here:
//   - a grid-wide barrier (cooperative groups), the form the stop decision
//     needs (every tile reads every tile's residual partial);
//   - flags with the four neighbours (release after the tile's stores,
//     acquire before the next step), the least a halo exchange needs.
// Across the chip's XCDs a release at agent scope writes back the XCD's L2
// and an acquire invalidates it (buffer_wbl2 / buffer_inv in the assembly),
// which is what these kernels time.  Every wait is bounded: a workgroup that
// polls a neighbour flag `spin_limit` times sets *err and goes on, so every
// wave reaches the end of the kernel.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace stm {

namespace {

constexpr int SYNC_TILE = 64 * 64;

__device__ __forceinline__ void sync_write_tile(double* __restrict__ out, int step) {
    double* o = out + (size_t)blockIdx.x * SYNC_TILE;
    for (int i = threadIdx.x; i < SYNC_TILE; i += SYNC_THREADS) o[i] = (double)(step + i);
}

__global__ void __launch_bounds__(SYNC_THREADS) k_sync_tile(double* __restrict__ out, int step) {
    sync_write_tile(out, step);
}

// holds every workgroup for `ticks` of the constant-rate wall clock
__global__ void __launch_bounds__(SYNC_THREADS) k_sync_hold(double* __restrict__ out, int64_t ticks) {
    const int64_t t0 = wall_clock64();
    while (wall_clock64() - t0 < ticks) {}
    if (threadIdx.x == 0) out[blockIdx.x] = 1.0;
}

__global__ void __launch_bounds__(SYNC_THREADS) k_sync_grid(double* __restrict__ out, int steps,
                                                            int write) {
    cooperative_groups::grid_group g = cooperative_groups::this_grid();
    for (int s = 0; s < steps; s++) {
        if (write) sync_write_tile(out, s);
        g.sync();
    }
}

__global__ void __launch_bounds__(SYNC_THREADS)
k_sync_neighbours(double* __restrict__ out, unsigned* flags, int side, int steps, int write,
                  unsigned spin_limit, unsigned* err) {
    const int b = blockIdx.x, bx = b / side, by = b - (b / side) * side;
    for (int s = 0; s < steps; s++) {
        if (write) sync_write_tile(out, s);
        __syncthreads();
        if (threadIdx.x == 0) {
            __threadfence();   // release: the tile before the flag
            __hip_atomic_store(&flags[b], (unsigned)(s + 1), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        if (threadIdx.x < 4) {
            const int nx = bx + (threadIdx.x == 0 ? -1 : threadIdx.x == 1 ? 1 : 0);
            const int ny = by + (threadIdx.x == 2 ? -1 : threadIdx.x == 3 ? 1 : 0);
            if (nx >= 0 && nx < side && ny >= 0 && ny < side) {
                unsigned n = 0;
                while (__hip_atomic_load(&flags[nx * side + ny], __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT) < (unsigned)(s + 1)) {
                    if (++n > spin_limit) {
                        atomicOr(err, 1u);
                        break;
                    }
                }
                __threadfence();   // acquire: the neighbour's tile after its flag
            }
        }
        __syncthreads();
    }
}

}  // namespace

void launch_sync_tile(hipStream_t s, int nwg, double* out, int step) {
    k_sync_tile<<<nwg, SYNC_THREADS, 0, s>>>(out, step);
}

void launch_sync_hold(hipStream_t s, int nwg, double* out, int64_t ticks) {
    k_sync_hold<<<nwg, SYNC_THREADS, 0, s>>>(out, ticks);
}

hipError_t sync_coresident(int nwg, int device) {
    int ncu = 0, a = 0, b = 0;
    hipError_t e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, (const void*)k_sync_grid,
                                                          SYNC_THREADS, 0);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, (const void*)k_sync_neighbours,
                                                          SYNC_THREADS, 0);
    if (e != hipSuccess) return e;
    return (ncu * a >= nwg && ncu * b >= nwg) ? hipSuccess : hipErrorCooperativeLaunchTooLarge;
}

hipError_t launch_sync_grid(hipStream_t s, int nwg, double* out, int steps, int write) {
    void* args[] = {&out, &steps, &write};
    return hipLaunchCooperativeKernel((const void*)k_sync_grid, dim3(nwg), dim3(SYNC_THREADS),
                                      args, 0, s);
}

hipError_t launch_sync_neighbours(hipStream_t s, int side, double* out, unsigned* flags,
                                  int steps, int write, unsigned spin_limit, unsigned* err) {
    void* args[] = {&out, &flags, &side, &steps, &write, &spin_limit, &err};
    return hipLaunchCooperativeKernel((const void*)k_sync_neighbours, dim3(side * side),
                                      dim3(SYNC_THREADS), args, 0, s);
}

}  // namespace stm
