/* artful-hip — the disc fluid's steps on the GPU, bit for bit the host's (scene/fluid.c, after Artful
 * 0.4.0's src/scene/fluid.c:96-241).
 *
 * The flow is chaotic, so every cell must see the very operations of the host code in the same order:
 * each cell's update is scene/fluid_cell.h's, the source the host's passes build from too, here built
 * with -ffp-contract=off (no fused multiply-add) and IEEE division. Each pass reads the previous
 * buffers and writes one cell per lane, as the host's rows do:
 *   - seed: the starting clumps (fluid.c:286-305): wrapped value-noise octaves, wound by the shear;
 *   - stir: a list of eddies applied in their order to each cell; a cell finds the one offset di of an
 *     eddy's sweep that lands on it, so it sees the same terms the host's loop over di gives it;
 *   - velocity, density: per cell;
 *   - shift and balance, a workgroup a ring: the ring turned by its orbit with a clamped Catmull-Rom
 *     cubic, its density kept in LDS, summed there by one lane in cell order (the host's order; a tree
 *     would round differently), then scaled by every lane. (As two passes, the shift writing the density
 *     and the balance's one lane reading it back from memory cell by cell, the balance alone took 23 us
 *     of a step's ~50.)
 *   - field: what shows.
 * Not in the reference (which steps its fluid on the CPU for every backend). */
#include <hip/hip_runtime.h>

#include "launch.h"
#include "scene/fluid_cell.h"

namespace artf {
namespace {

constexpr int FW = FLUID_W, FH = FLUID_H;

/* The cell a lane owns: lane number n is cell n % 1024 of ring n / 1024. */
__device__ __forceinline__ bool cellOf(int& i, int& j) {
    const int at = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    i = at & (FW - 1), j = at / FW;
    return j < FH;
}

__global__ __launch_bounds__(256) void seedKernel(ArtfHipFluid f) {
    int i, j;
    if (!cellOf(i, j)) return;
    const size_t at = (size_t)j * FW + i;
    f.density[at] = fluidSeedCell(f.rings, i, j);
    f.radial[at] = 0.f, f.around[at] = 0.f;
    f.scratch[1][at] = 0.f, f.scratch[2][at] = 0.f;
}

__global__ __launch_bounds__(256) void stirKernel(ArtfHipFluid f, const ArtfFluidEddy* __restrict__ eddies, int count) {
    int i, j;
    if (!cellOf(i, j)) return;
    const float rj = f.rings->r[j];
    const size_t at = (size_t)j * FW + i;
    float s = f.density[at], vr = f.radial[at], w = f.around[at];
    for (int n = 0; n < count; ++n) {
        const ArtfFluidEddy e = eddies[n];
        const float dr = rj - e.r0;
        if (dr * dr > 6.25f * e.size * e.size) continue;
        const int di = ((i - e.i0 + FW / 2) & (FW - 1)) - FW / 2; /* the sweep's offset that lands here */
        if (di < -e.reach || di > e.reach) continue;
        fluidEddyCell(&e, dr, di, rj, &s, &vr, &w);
    }
    f.density[at] = s, f.radial[at] = vr, f.around[at] = w;
}

__global__ __launch_bounds__(256) void velocityKernel(ArtfHipFluid f) {
    int i, j;
    if (!cellOf(i, j)) return;
    const FluidRow q = fluidRowOf(f.rings, j);
    const size_t at = (size_t)j * FW + i;
    fluidVelocityCell(&q, f.density, f.radial, f.around, i, j, f.scratch[1] + at, f.scratch[2] + at);
}

__global__ __launch_bounds__(256) void densityKernel(ArtfHipFluid f) {
    int i, j;
    if (!cellOf(i, j)) return;
    const FluidRow q = fluidRowOf(f.rings, j);
    f.scratch[0][(size_t)j * FW + i] = fluidDensityCell(&q, f.density, f.scratch[1], f.scratch[2], i, j);
}

/* A workgroup a ring (fluid.c:153-161, 221-229): the ring turned along itself by its orbit's cells a
 * step, from the step's second buffers back into the state, its density through LDS; one lane sums the
 * turned density in cell order, then every lane scales four cells by the ring's balance. */
__global__ __launch_bounds__(256) void ringKernel(ArtfHipFluid f) {
    __shared__ __align__(16) float turned[FW];
    __shared__ float scale;
    const int j = (int)blockIdx.x;
    const float cells = f.rings->shiftCells[j];
    const size_t row = (size_t)j * FW;
    for (int i = (int)threadIdx.x; i < FW; i += (int)blockDim.x) {
        turned[i] = fluidShiftCell(f.scratch[0] + row, i, cells);
        f.radial[row + i] = fluidShiftCell(f.scratch[1] + row, i, cells);
        f.around[row + i] = fluidShiftCell(f.scratch[2] + row, i, cells);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        /* the chain of additions in cell order, the LDS reads four cells at a time and issued ahead of it */
        const float4* quads = reinterpret_cast<const float4*>(turned);
        float sum = 0;
#pragma unroll 16
        for (int i = 0; i < FW / 4; ++i) {
            const float4 v = quads[i];
            sum += v.x, sum += v.y, sum += v.z, sum += v.w;
        }
        scale = fluidBalanceScale(sum);
    }
    __syncthreads();
    for (int i = (int)threadIdx.x; i < FW; i += (int)blockDim.x) f.density[row + i] = sharedClamp(turned[i] * scale, .03f, 8.f);
}

__global__ __launch_bounds__(256) void fieldKernel(ArtfHipFluid f) {
    int i, j;
    if (!cellOf(i, j)) return;
    const size_t at = (size_t)j * FW + i;
    f.field[at] = fluidFieldCell(f.density[at], f.rings->envelope[j], f.rings->falloff[j]);
}

constexpr unsigned kCellBlocks = (unsigned)(FW * FH / 256);

} // namespace

hipError_t launchFluidSeed(const ArtfHipFluid& f, hipStream_t stream) {
    hipLaunchKernelGGL(seedKernel, dim3(kCellBlocks), dim3(256), 0, stream, f);
    return hipGetLastError();
}

hipError_t launchFluidStir(const ArtfHipFluid& f, const ArtfFluidEddy* eddies, int count, hipStream_t stream) {
    hipLaunchKernelGGL(stirKernel, dim3(kCellBlocks), dim3(256), 0, stream, f, eddies, count);
    return hipGetLastError();
}

hipError_t launchFluidStep(const ArtfHipFluid& f, const ArtfFluidEddy* eddies, hipStream_t stream) {
    hipLaunchKernelGGL(stirKernel, dim3(kCellBlocks), dim3(256), 0, stream, f, eddies, ARTF_FLUID_STEP_EDDIES);
    hipLaunchKernelGGL(velocityKernel, dim3(kCellBlocks), dim3(256), 0, stream, f);
    hipLaunchKernelGGL(densityKernel, dim3(kCellBlocks), dim3(256), 0, stream, f);
    hipLaunchKernelGGL(ringKernel, dim3((unsigned)FH), dim3(256), 0, stream, f);
    return hipGetLastError();
}

hipError_t launchFluidField(const ArtfHipFluid& f, hipStream_t stream) {
    hipLaunchKernelGGL(fieldKernel, dim3(kCellBlocks), dim3(256), 0, stream, f);
    return hipGetLastError();
}

} // namespace artf
