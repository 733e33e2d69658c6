/* artful-hip — the disc map's painted tables on the GPU, bit for bit the host's (scene/disc.c, after
 * Artful 0.4.0's src/scene/hole.c:10-178): the gas (2048 x 1024), normalised to a mean of 1 over its
 * inner half; the disc's dust lanes (1024 x 256); the sky's dust (1024 x 512). The colour table and the
 * ring tables are small and stay on the host.
 *
 * Built with -ffp-contract=off and IEEE division: each texel is scene/disc_texel.h's, the very source
 * the host paints with, from its coordinates alone (a lattice cell below zero goes through a signed
 * conversion there, as on x86). One place needs care of its own to give the host's bits:
 *   - the normalising sum over 1,048,576 texels is taken in doubles in texel order, as the host's
 *     loop does: one lane adds them in order from LDS while the block's other waves stage the next
 *     4096 (a tree would round differently).
 * Not in the reference (which paints on the CPU, 0.6 s). */
#include <hip/hip_runtime.h>

#include "launch.h"
#include "scene/disc_texel.h"

namespace artf {
namespace {

constexpr int MW = ARTF_HOLE_MAP_WIDTH, MH = ARTF_HOLE_MAP_HEIGHT;
constexpr int DW = ARTF_HOLE_DUST_WIDTH, DH = ARTF_HOLE_DUST_HEIGHT;
constexpr int SW = ARTF_HOLE_SKY_WIDTH, SH = ARTF_HOLE_SKY_HEIGHT;

__global__ __launch_bounds__(256) void gasKernel(float* map) {
    const int at = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (at >= MW * MH) return;
    map[at] = discGasTexel(at % MW, at / MW);
}

/* The sum of the inner half's texels in doubles, in texel order (disc.c artfDiscPaint), and the gas
 * scaled to make their mean 1. One block: lane 0 adds each run of texels from LDS in order, four at a
 * read issued ahead of the chain, while the other three waves stage the next run in the other half of
 * LDS; then every lane scales. */
constexpr int kRun = 4096;

__global__ __launch_bounds__(256) void gasScaleKernel(float* map) {
    __shared__ __align__(16) float run[2][kRun];
    __shared__ float scale;
    const int inner = MW * (MH / 2), t = (int)threadIdx.x;
    for (int i = t; i < kRun; i += (int)blockDim.x) run[0][i] = map[i];
    __syncthreads();
    double sum = 0;
    for (int base = 0, b = 0; base < inner; base += kRun, b ^= 1) {
        if (t == 0) {
            const float4* quads = reinterpret_cast<const float4*>(run[b]);
#pragma unroll 8
            for (int i = 0; i < kRun / 4; ++i) {
                const float4 v = quads[i];
                sum += v.x, sum += v.y, sum += v.z, sum += v.w;
            }
        } else if (t >= 64 && base + kRun < inner) {
            for (int i = t - 64; i < kRun; i += (int)blockDim.x - 64) run[b ^ 1][i] = map[base + kRun + i];
        }
        __syncthreads();
    }
    if (t == 0) scale = (float)((double)inner / sum);
    __syncthreads();
    const float k = scale;
    for (int i = t; i < MW * MH; i += (int)blockDim.x) map[i] *= k;
}

__global__ __launch_bounds__(256) void dustKernel(float* dust) {
    const int at = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (at >= DW * DH) return;
    dust[at] = discDustTexel(at % DW, at / DW);
}

__global__ __launch_bounds__(256) void skyKernel(float* sky) {
    const int at = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (at >= SW * SH) return;
    sky[at] = discSkyTexel(at % SW, at / SW);
}

unsigned blocksFor(int n) { return (unsigned)((n + 255) / 256); }

} // namespace

hipError_t launchDiscPaint(float* map, hipStream_t stream) {
    hipLaunchKernelGGL(gasKernel, dim3(blocksFor(MW * MH)), dim3(256), 0, stream, map);
    hipLaunchKernelGGL(gasScaleKernel, dim3(1), dim3(256), 0, stream, map);
    hipLaunchKernelGGL(dustKernel, dim3(blocksFor(DW * DH)), dim3(256), 0, stream, map + ARTF_HOLE_DUST_AT);
    hipLaunchKernelGGL(skyKernel, dim3(blocksFor(SW * SH)), dim3(256), 0, stream, map + ARTF_HOLE_SKY_AT);
    return hipGetLastError();
}

} // namespace artf
