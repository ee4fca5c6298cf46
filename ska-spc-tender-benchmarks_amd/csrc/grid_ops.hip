// Adder (uv subgrids onto the master grid) and splitter (master grid into uv subgrids) for gfx950.
//
// Grid layout: complex64 [4][G][G] (one plane per correlation).  Subgrid s covers grid cells
// [y0, y0 + S) x [x0, x0 + S) with (x0, y0) = metadata[s].coordinate; cells outside the grid are skipped by the
// adder and read as zero by the splitter.
//
// Conflict-free adder.  The grid is cut into tiles of T >= S cells; a subgrid belongs to the tile holding its
// corner and can only touch that tile and its right/lower neighbours.  Tiles are coloured by the parity of their
// tile coordinates: two different tiles of one colour are at least two tiles apart in x or y, so their subgrids
// never touch the same cell.  One launch per colour; one workgroup per tile adds that tile's subgrids one after
// another (a barrier between subgrids), so every grid cell sees a fixed summation order: the result is
// deterministic and needs no atomics.  The tile lists are built on the host from the plan
// (idgx_adder_plan_build).
#include "idgx_device.hpp"
#include "kernels.hpp"

namespace idgx
{
namespace
{
constexpr int kThreads = 256;

__global__ __launch_bounds__(kThreads) void adder_kernel(const idgx_metadata *md, const int32_t *order,
                                                         const int32_t *tile_start, int tile_begin, int S, int G,
                                                         const float2 *subgrids, float2 *grid)
{
    const int tile = tile_begin + blockIdx.x;
    const int first = tile_start[tile], last = tile_start[tile + 1];
    const int P = S * S;
    const size_t GG = static_cast<size_t>(G) * G;
    for (int i = first; i < last; ++i)
    {
        const int s = order[i];
        const int x0 = md[s].x, y0 = md[s].y;
        const float2 *sub = subgrids + static_cast<size_t>(s) * kCorr * P;
        for (int q = threadIdx.x; q < kCorr * P; q += kThreads)
        {
            const int pol = q / P, rem = q - pol * P;
            const int y = rem / S, x = rem - y * S;
            const int gy = y0 + y, gx = x0 + x;
            if (gy >= 0 && gy < G && gx >= 0 && gx < G)
            {
                float2 *g = grid + pol * GG + static_cast<size_t>(gy) * G + gx;
                const float2 v = sub[q];
                const float2 o = *g;
                *g = make_float2(o.x + v.x, o.y + v.y);
            }
        }
        __syncthreads(); // the next subgrid of this tile may overlap this one
    }
}

__global__ __launch_bounds__(kThreads) void splitter_kernel(const idgx_metadata *md, int S, int G, const float2 *grid,
                                                            float2 *subgrids)
{
    const long long s = blockIdx.x;
    const int x0 = md[s].x, y0 = md[s].y;
    const int P = S * S;
    const size_t GG = static_cast<size_t>(G) * G;
    float2 *sub = subgrids + static_cast<size_t>(s) * kCorr * P;
    for (int q = threadIdx.x; q < kCorr * P; q += kThreads)
    {
        const int pol = q / P, rem = q - pol * P;
        const int y = rem / S, x = rem - y * S;
        const int gy = y0 + y, gx = x0 + x;
        float2 v = make_float2(0.F, 0.F);
        if (gy >= 0 && gy < G && gx >= 0 && gx < G)
            v = grid[pol * GG + static_cast<size_t>(gy) * G + gx];
        sub[q] = v;
    }
}
} // namespace

hipError_t launch_adder(const idgx_metadata *md, const int32_t *order, const int32_t *tile_start,
                        const int32_t *color_begin, int S, int G, const void *subgrids, void *grid, hipStream_t st)
{
    for (int c = 0; c < 4; ++c)
    {
        const int n = color_begin[c + 1] - color_begin[c];
        if (n <= 0)
            continue;
        adder_kernel<<<n, kThreads, 0, st>>>(md, order, tile_start, color_begin[c], S, G,
                                             static_cast<const float2 *>(subgrids), static_cast<float2 *>(grid));
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

hipError_t launch_splitter(const idgx_metadata *md, long long nr_subgrids, int S, int G, const void *grid,
                           void *subgrids, hipStream_t st)
{
    if (nr_subgrids > 0x7fffffffLL)
        return hipErrorInvalidConfiguration;
    splitter_kernel<<<static_cast<unsigned>(nr_subgrids), kThreads, 0, st>>>(
        md, S, G, static_cast<const float2 *>(grid), static_cast<float2 *>(subgrids));
    return hipGetLastError();
}
} // namespace idgx
