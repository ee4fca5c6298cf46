// Adder (uv subgrids onto the master grid) and splitter (master grid into uv subgrids) for gfx950.
//
// Grid layout: complex64 [4][G][G] (one plane per correlation), or [nr_layers][4][G][G] for w-stacking, where
// subgrid s belongs to layer metadata[s].z - z_min.  Subgrid s covers grid cells
// [y0, y0 + S) x [x0, x0 + S) with (x0, y0) = metadata[s].coordinate; cells outside the grid are skipped by the
// adder and read as zero by the splitter.
//
// Conflict-free adder.  The grid is cut into tiles of T >= S cells; a subgrid belongs to the tile holding its
// corner and can only touch that tile's 2T x 2T region (the tile and its right/lower neighbours).  Tiles are
// coloured by the parity of their tile coordinates: two different tiles of one colour are at least two tiles
// apart in x or y, so their regions are disjoint.  One launch per colour.  A workgroup takes one correlation of
// one row block (at most 1024 cells) of one quadrant of a tile's region, walks the tile's subgrids in plan order
// accumulating the covered cells in registers, and then adds each accumulated cell to the grid once.  Every
// cell therefore sees a fixed summation order (deterministic, no atomics), and a tile crowded with subgrids is
// spread over 16 x (row blocks) workgroups instead of one.  The tile lists are built on the host from the plan
// (idgx_adder_plan_build).
#include "idgx_device.hpp"
#include "kernels.hpp"

namespace idgx
{
namespace
{
constexpr int kThreads = 256;

constexpr int kAdderCells = 4;  // grid cells per thread of the adder
constexpr int kAdderUnroll = 4; // subgrids whose loads the adder keeps in flight together

__device__ __forceinline__ int floordiv_dev(int a, int b)
{
    return a >= 0 ? a / b : -((-a + b - 1) / b);
}

// Rows of a tile quadrant one adder workgroup covers (at most kThreads * kAdderCells cells).
__host__ __device__ inline int adder_rows(int T)
{
    const int r = (kThreads * kAdderCells) / T;
    return r < T ? r : T;
}

__global__ __launch_bounds__(kThreads) void adder_kernel(const idgx_metadata *md, const int32_t *order,
                                                         const int32_t *tile_start, int tile_begin, int ntiles, int S,
                                                         int T, int G, int z_min, int nr_layers,
                                                         const float2 *subgrids, float2 *grid)
{
    const int R = adder_rows(T), nrb = (T + R - 1) / R;
    const int per_tile = 16 * nrb;
    // Workgroup b -> (tile, sub-block): workgroups reach the 8 XCDs in turn (b % 8), so XCD x takes tiles x, x + 8,
    // x + 16, ... of the colour (heaviest first), each with all its sub-blocks.  The rows of a subgrid that straddle
    // the boundary between two quadrants are then read by workgroups of one XCD, so each cache line they share is
    // fetched into one L2 once.  The grid is padded to a multiple of 8 tiles; the padding exits.
    const int b = static_cast<int>(blockIdx.x), k = b >> 3;
    const int tl = (k / per_tile) * 8 + (b & 7);
    if (tl >= ntiles)
        return; // uniform
    const int tile = tile_begin + tl;
    const int sub_b = k % per_tile;
    const int pol = sub_b / (4 * nrb), quad = (sub_b / nrb) & 3, rb = sub_b % nrb;
    const int first = tile_start[tile], last = tile_start[tile + 1];
    const idgx_metadata m0 = md[order[first]];
    // All subgrids of a tile share one w-layer (the plan keys tiles by layer); single-layer plans ignore z.
    const int layer = nr_layers > 1 ? m0.z - z_min : 0;
    if (layer < 0 || layer >= nr_layers)
        return; // uniform; the layered plan builder rejects such subgrids
    // Cells [by, by + nrows) x [bx, bx + T) of the grid.  The tile origin is taken from the tile index as the plan
    // keys it (clamped to -1 .. G/T + 1), so every subgrid of the tile agrees on it.
    const int ntiles_axis = floordiv_dev(G, T) + 3;
    const int tx = min(max(floordiv_dev(m0.x, T) + 1, 0), ntiles_axis - 1) - 1;
    const int ty = min(max(floordiv_dev(m0.y, T) + 1, 0), ntiles_axis - 1) - 1;
    const int bx = tx * T + (quad & 1) * T;
    const int by = ty * T + (quad >> 1) * T + rb * R;
    const int nrows = min(R, T - rb * R);
    const int ncells = nrows * T;
    const int P = S * S;
    float2 acc[kAdderCells];
    bool touched[kAdderCells];
    int cx[kAdderCells], cy[kAdderCells];
#pragma unroll
    for (int k = 0; k < kAdderCells; ++k)
    {
        const int c = threadIdx.x + k * kThreads;
        cy[k] = by + c / T;
        cx[k] = bx + c % T;
        acc[k] = make_float2(0.F, 0.F);
        touched[k] = false;
        if (c >= ncells)
            cy[k] = -0x40000000; // never covered
    }
    // kAdderUnroll subgrids per step: their loads are all in flight together, then added in plan order.
    for (int i0 = first; i0 < last; i0 += kAdderUnroll)
    {
        float2 val[kAdderUnroll][kAdderCells];
        bool hit[kAdderUnroll][kAdderCells];
#pragma unroll
        for (int u = 0; u < kAdderUnroll; ++u)
        {
            const int i = i0 + u;
            int s = 0, x0 = 0, y0 = 0;
            bool any = i < last;
            if (any)
            {
                s = order[i];
                x0 = md[s].x;
                y0 = md[s].y;
                any = !(x0 + S <= bx || x0 >= bx + T || y0 + S <= by || y0 >= by + nrows); // uniform
            }
            const float2 *sub = subgrids + (static_cast<size_t>(s) * kCorr + pol) * P;
#pragma unroll
            for (int k = 0; k < kAdderCells; ++k)
            {
                const int dx = cx[k] - x0, dy = cy[k] - y0;
                hit[u][k] = any && dx >= 0 && dx < S && dy >= 0 && dy < S;
                val[u][k] = hit[u][k] ? sub[dy * S + dx] : make_float2(0.F, 0.F);
            }
        }
#pragma unroll
        for (int u = 0; u < kAdderUnroll; ++u)
#pragma unroll
            for (int k = 0; k < kAdderCells; ++k)
                if (hit[u][k])
                {
                    acc[k].x += val[u][k].x;
                    acc[k].y += val[u][k].y;
                    touched[k] = true;
                }
    }
    const size_t GG = static_cast<size_t>(G) * G;
    float2 *plane = grid + (static_cast<size_t>(layer) * kCorr + pol) * GG;
#pragma unroll
    for (int k = 0; k < kAdderCells; ++k)
    {
        if (touched[k] && cx[k] >= 0 && cx[k] < G && cy[k] >= 0 && cy[k] < G)
        {
            float2 *g = plane + static_cast<size_t>(cy[k]) * G + cx[k];
            const float2 o = *g;
            *g = make_float2(o.x + acc[k].x, o.y + acc[k].y);
        }
    }
}

__global__ __launch_bounds__(kThreads) void splitter_kernel(const idgx_metadata *md, int S, int G, int z_min,
                                                            int nr_layers, const float2 *grid, float2 *subgrids)
{
    // Subgrids in launch order, spread over the XCDs: an XCD-contiguous order (neighbours in uv on one XCD) read
    // 1.52 GB instead of 3.13 GB for 76,865 subgrids of 32^2 on a 4096^2 grid, but took 0.98 ms instead of 0.84 ms.
    const long long s = blockIdx.x;
    const int x0 = md[s].x, y0 = md[s].y;
    const int P = S * S;
    const size_t GG = static_cast<size_t>(G) * G;
    const int layer = nr_layers > 1 ? md[s].z - z_min : 0;
    const bool on = layer >= 0 && layer < nr_layers; // uniform: a subgrid of no layer reads zeros
    grid += static_cast<size_t>(on ? layer : 0) * kCorr * GG;
    float2 *sub = subgrids + static_cast<size_t>(s) * kCorr * P;
    for (int q = threadIdx.x; q < kCorr * P; q += kThreads)
    {
        const int pol = q / P, rem = q - pol * P;
        const int y = rem / S, x = rem - y * S;
        const int gy = y0 + y, gx = x0 + x;
        float2 v = make_float2(0.F, 0.F);
        if (on && gy >= 0 && gy < G && gx >= 0 && gx < G)
            v = grid[pol * GG + static_cast<size_t>(gy) * G + gx];
        sub[q] = v;
    }
}
} // namespace

hipError_t launch_adder(const idgx_metadata *md, const int32_t *order, const int32_t *tile_start,
                        const int32_t *color_begin, int S, int T, int G, int z_min, int nr_layers,
                        const void *subgrids, void *grid, hipStream_t st)
{
    if (T < S || T <= 0)
        return hipErrorInvalidValue;
    const int R = adder_rows(T);
    const long long per_tile = 16LL * ((T + R - 1) / R);
    for (int c = 0; c < 4; ++c)
    {
        const long long n = color_begin[c + 1] - color_begin[c];
        if (n <= 0)
            continue;
        const long long padded = (n + 7) / 8 * 8; // whole groups of 8 tiles, one per XCD
        if (padded * per_tile > 0x7fffffffLL)
            return hipErrorInvalidConfiguration;
        adder_kernel<<<static_cast<unsigned>(padded * per_tile), kThreads, 0, st>>>(
            md, order, tile_start, color_begin[c], static_cast<int>(n), S, T, G, z_min, nr_layers,
            static_cast<const float2 *>(subgrids), static_cast<float2 *>(grid));
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

hipError_t launch_splitter(const idgx_metadata *md, long long nr_subgrids, int S, int G, int z_min, int nr_layers,
                           const void *grid, void *subgrids, hipStream_t st)
{
    if (nr_subgrids > 0x7fffffffLL)
        return hipErrorInvalidConfiguration;
    splitter_kernel<<<static_cast<unsigned>(nr_subgrids), kThreads, 0, st>>>(
        md, S, G, z_min, nr_layers, static_cast<const float2 *>(grid), static_cast<float2 *>(subgrids));
    return hipGetLastError();
}
} // namespace idgx
