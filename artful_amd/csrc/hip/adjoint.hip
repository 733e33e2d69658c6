/* artful-hip — the adjoint kernel: each lane runs one ray forward and back (adjoint.h), and each
 * workgroup (one wave over an 8 x 8 tile, as the tile tracer) sums its rays' gradients into 32 doubles
 * (a fixed order: the sums are the same every run). Built without contraction, so its forward light
 * equals trace_exact.hip's analytic-disc light (the reverse pass's own arithmetic fuses: adjoint.h). */
#include "adjoint.h"
#include "trace_kernel.h"

namespace artf {

/* Waves a SIMD the adjoint's registers are held to: the FP32 one at 3 (168 VGPRs; its next-step loads
 * ahead take 175 and 2 waves otherwise, 3% slower); the FP64 one keeps what it needs (1 wave). */
constexpr int adjointWavesPerSimd(int bytes) { return bytes == 4 ? 3 : 1; }

template <typename R>
__global__ __launch_bounds__(kTraceGroup) __attribute__((amdgpu_waves_per_eu(adjointWavesPerSimd(sizeof(R))))) void adjointKernel(ArtfHipAdjointArgs args) {
    const ArtfHipTraceArgs& a = args.trace;
    const ArtfHoleFrame& f = a.frame;
    const int width = (int)f.size[0], height = (int)f.size[1], side = (int)f.size[2];
    const int lane = (int)threadIdx.x;
    const int px = (int)blockIdx.x * kTraceTile + (lane & 7);
    const int py = (int)blockIdx.y * kTraceTile + (lane >> 3);
    const int sub = blockIdx.z, sx = sub % side, sy = sub / side;
    const bool valid = px < width && py < height;

    TraceLaw law;
    law.maxSteps = a.maxSteps > 0 ? a.maxSteps : 192;
    law.uEscape = a.escapeRadius > 0.f ? 1.f / a.escapeRadius : 0.f;
    law.disc = DISC_ANALYTIC;
    law.q = a.emissivity;
    law.inRangeDrag = false; /* the adjoint's rays take no drag (adjointRay) */

    double grad[ADJ_COUNT];
    for (int k = 0; k < ADJ_COUNT; ++k) grad[k] = 0.0;
    if (valid) {
        float dx, dy;
        subPixel(a.jitter, px, py, sx, sy, side, dx, dy);
        const float moment = ((float)((sub * 5) % (side * side)) + .5f) * f.image[3] - .5f;
        const float time = f.disc[2] + f.spin[3] * moment;
        const size_t pixel = (size_t)py * width + px;
        float light[3];
        if (sizeof(R) == 8) {
            Eye<double> e;
            eyeFrom64(a.frame64, e);
            const double nx = ((double)px + (double)dx) * (2.0 / (double)width) - 1.0;
            const double ny = 1.0 - ((double)py + (double)dy) * (2.0 / (double)height);
            const double seed[3] = {(double)args.seed[0][pixel] * f.image[3], (double)args.seed[1][pixel] * f.image[3],
                                    (double)args.seed[2][pixel] * f.image[3]};
            Saved<double> ckpt[ARTF_ADJ_SLOTS], seg[ARTF_ADJ_CHECKPOINT + 1];
            double g[ADJ_COUNT];
            for (int k = 0; k < ADJ_COUNT; ++k) g[k] = 0.0;
            adjointRay<double>(f, e, law, a.map, nx, ny, time, seed, light, g, ckpt, seg);
            for (int k = 0; k < ADJ_COUNT; ++k) grad[k] = g[k];
        } else {
            Eye<float> e;
            eyeFrom<float>(f, e);
            const float nx = ((float)(px & 7) + ((float)(px & ~7) + dx)) * f.image[1] - 1.f;
            const float ny = 1.f - ((float)py + dy) * f.image[2];
            const float seed[3] = {args.seed[0][pixel] * f.image[3], args.seed[1][pixel] * f.image[3], args.seed[2][pixel] * f.image[3]};
            Saved<float> ckpt[ARTF_ADJ_SLOTS], seg[ARTF_ADJ_CHECKPOINT + 1];
            float g[ADJ_COUNT];
            for (int k = 0; k < ADJ_COUNT; ++k) g[k] = 0.f;
            adjointRay<float>(f, e, law, a.map, nx, ny, time, seed, light, g, ckpt, seg);
            for (int k = 0; k < ADJ_COUNT; ++k) grad[k] = (double)g[k];
        }
        const size_t at = (side == 1 ? 0 : (size_t)sub * width * height) + pixel;
        a.out[0][at] = light[0];
        a.out[1][at] = light[1];
        a.out[2][at] = light[2];
    }
    /* the wave's sum, in a fixed order (the same butterfly on every lane), stored by lane k */
    const size_t block = ((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    double mine = 0.0;
    for (int k = 0; k < ADJ_COUNT; ++k) {
        double v = grad[k];
        for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_xor(v, offset);
        if (lane == k) mine = v;
    }
    if (lane < 32) args.partials[block * 32 + lane] = mine;
}

int adjointBlocks(const ArtfHoleFrame& f) {
    const int side = (int)f.size[2];
    return (int)(((f.size[0] + kTraceTile - 1) / kTraceTile) * ((f.size[1] + kTraceTile - 1) / kTraceTile)) * side * side;
}

hipError_t launchAdjoint(const ArtfHipAdjointArgs& args, hipStream_t stream) {
    const ArtfHoleFrame& f = args.trace.frame;
    const int width = (int)f.size[0], height = (int)f.size[1], side = (int)f.size[2];
    const dim3 grid((unsigned)((width + kTraceTile - 1) / kTraceTile), (unsigned)((height + kTraceTile - 1) / kTraceTile), (unsigned)(side * side));
    if (args.trace.precision == ARTF_PRECISION_FP64) hipLaunchKernelGGL(adjointKernel<double>, grid, dim3(kTraceGroup), 0, stream, args);
    else hipLaunchKernelGGL(adjointKernel<float>, grid, dim3(kTraceGroup), 0, stream, args);
    return hipGetLastError();
}

} // namespace artf
