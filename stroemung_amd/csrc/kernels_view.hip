// gfx950 kernel of the GUI image (src/visualization.rs:7-105): every cell of
// a channel coloured by pressure or speed on the device, so a viewer reads
// back 4 bytes per cell instead of three FP64 fields.  The arithmetic follows
// the reference: the hue in FP64 cast to f32, the HSL conversion in f32, and
// macroquad 0.4.13's Color -> [u8; 4] (component * 255 truncated, saturating,
// NaN -> 0).  Built with -ffp-contract=off like every file here, so the f32
// products are not fused and the image is bit-identical to the reference's.
#include <hip/hip_runtime.h>

#include <mutex>

#include "kernels.h"
#include "stm_internal.h"

namespace stm {

namespace {

__device__ __forceinline__ void hsl_to_rgb(float hue, float sat, float light, float& r, float& g,
                                           float& b) {
    const float c = (1.0f - fabsf(2.0f * light - 1.0f)) * sat;
    const float x = c * (1.0f - fabsf(fmodf(hue / 60.0f, 2.0f) - 1.0f));
    const float m = light - c / 2.0f;
    float rr, gg, bb;
    if (hue < 60.0f) { rr = c; gg = x; bb = 0.0f; }
    else if (hue < 120.0f) { rr = x; gg = c; bb = 0.0f; }
    else if (hue < 180.0f) { rr = 0.0f; gg = c; bb = x; }
    else if (hue < 240.0f) { rr = 0.0f; gg = x; bb = c; }
    else if (hue < 300.0f) { rr = x; gg = 0.0f; bb = c; }
    else { rr = c; gg = 0.0f; bb = x; }
    r = rr + m;
    g = gg + m;
    b = bb + m;
}

__device__ __forceinline__ uint32_t to_u8(float c) {
    const float v = c * 255.0f;
    if (!(v > 0.0f)) return 0u;
    if (v >= 255.0f) return 255u;
    return (uint32_t)v;
}

}  // namespace

// One thread per pixel; pixel (x, y) of channel `ch` at (y * nx + x) * 4.
__global__ void __launch_bounds__(256)
k_render(const ChanDesc* __restrict__ cd, const ChanState* __restrict__ cs, int ch,
         const uint8_t* __restrict__ flags, const double* __restrict__ u,
         const double* __restrict__ v, const double* __restrict__ p, int mode,
         uint32_t* __restrict__ rgba)
{
    const ChanDesc& d = cd[ch];
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)d.nx * d.ny) return;
    const int y = (int)(k / d.nx), x = (int)(k - (int64_t)y * d.nx);
    const int64_t i = d.off + x * (int64_t)d.pitch + y;
    float r, g, b;
    if (!flag_fluid(flags[i])) {
        r = 0.5f;
        g = mode ? 0.5f : 0.0f;
        b = mode ? 0.5f : 0.0f;
    } else {
        const ChanState& s = cs[ch];
        double t;
        if (mode) {
            const double speed = sqrt(u[i] * u[i] + v[i] * v[i]);
            t = 240.0 - (speed - s.speed_range[0]) * 240.0 / (s.speed_range[1] - s.speed_range[0]);
        } else {
            t = 240.0 - (p[i] - s.pressure_range[0]) * 240.0 /
                            (s.pressure_range[1] - s.pressure_range[0]);
        }
        hsl_to_rgb((float)t, 1.0f, 0.5f, r, g, b);
    }
    rgba[k] = to_u8(r) | (to_u8(g) << 8) | (to_u8(b) << 16) | (255u << 24);
}

// ---------------------------------------------------------------- static image buffer
//
// stm_render_static renders into this array instead of a buffer the library
// allocates.  Every device the library runs on holds its own copy until the
// process ends; host code reaches it only through the device address looked
// up once per device (the array's name on the host is a stand-in, not a
// device address).  Calls that overlap on one device share it and race.
static __device__ __align__(256) uint32_t scratch[RENDER_STATIC_PIXELS];

namespace {
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[RENDER_STATIC_DEVICES] = {};
}  // namespace

hipError_t render_static_init(int device) {
    if (device < 0 || device >= RENDER_STATIC_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));   // device is current
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = static_cast<uint32_t*>(ptr);
    return hipSuccess;
}

uint32_t* render_static_buffer(int device) {
    if (device < 0 || device >= RENDER_STATIC_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

void launch_render(hipStream_t st, const ChanDesc* cd, const ChanState* cs, int ch, int64_t npix,
                   const uint8_t* flags, const double* u, const double* v, const double* p,
                   int mode, uint32_t* rgba) {
    k_render<<<(unsigned)((npix + 255) / 256), 256, 0, st>>>(cd, cs, ch, flags, u, v, p, mode,
                                                             rgba);
}

}  // namespace stm
