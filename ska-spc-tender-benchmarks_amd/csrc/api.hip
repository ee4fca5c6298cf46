// C ABI of libidgx: argument validation and dispatch to the gfx950 kernels.
#include <algorithm>
#include <utility>
#include <vector>

#include "idgx.h"
#include "kernels.hpp"

namespace
{
using idgx::KernelArgs;
using idgx::PlanDev;

constexpr int kMinSubgrid = 8;
constexpr int kMaxSubgrid = 128;

int check_plan_struct(const idgx_plan *p)
{
    if (!p)
        return IDGX_ERR_INVALID_ARGUMENT;
    if (p->nr_subgrids < 0 || p->nr_channels <= 0 || p->nr_stations <= 0 || p->nr_aterm_slots <= 0 ||
        p->nr_rows < 0 || p->max_nr_timesteps < 0 || p->grid_size <= 0 || !(p->image_size > 0.F))
        return IDGX_ERR_INVALID_ARGUMENT;
    if (p->subgrid_size < kMinSubgrid || p->subgrid_size > kMaxSubgrid || (p->subgrid_size & 1))
        return IDGX_ERR_UNSUPPORTED;
    if (p->aterm_convention != IDGX_ATERM_ADJOINT && p->aterm_convention != IDGX_ATERM_REFERENCE)
        return IDGX_ERR_INVALID_ARGUMENT;
    if (p->precision < IDGX_PLAIN_FP32 || p->precision > IDGX_MIXED_FP16)
        return IDGX_ERR_INVALID_ARGUMENT;
    if (p->flags & ~(IDGX_FLAG_ACCURATE_SINCOS | IDGX_FLAG_PHASE_MFMA))
        return IDGX_ERR_INVALID_ARGUMENT;
    return IDGX_OK;
}

PlanDev to_dev(const idgx_plan *p)
{
    PlanDev d;
    d.nr_subgrids = p->nr_subgrids;
    d.S = p->subgrid_size;
    d.G = p->grid_size;
    d.nr_channels = p->nr_channels;
    d.nr_stations = p->nr_stations;
    d.nr_aterm_slots = p->nr_aterm_slots;
    d.max_nr_timesteps = p->max_nr_timesteps;
    d.nr_rows = p->nr_rows;
    d.image_size = p->image_size;
    d.w_step = p->w_step_in_lambda;
    d.convention = p->aterm_convention;
    return d;
}

bool accurate_sincos(const idgx_plan *p)
{
    return (p->flags & IDGX_FLAG_ACCURATE_SINCOS) != 0;
}

int to_status(hipError_t e)
{
    if (e == hipSuccess)
        return IDGX_OK;
    if (e == hipErrorOutOfMemory)
        return IDGX_ERR_OUT_OF_MEMORY;
    if (e == hipErrorInvalidConfiguration)
        return IDGX_ERR_UNSUPPORTED;
    return IDGX_ERR_DEVICE;
}
} // namespace

extern "C" int idgx_check_plan(const idgx_plan *plan, const idgx_metadata *md, const idgx_channels *ch)
{
    const int rc = check_plan_struct(plan);
    if (rc != IDGX_OK)
        return rc;
    if (plan->nr_subgrids > 0 && !md)
        return IDGX_ERR_INVALID_ARGUMENT;
    for (int s = 0; s < plan->nr_subgrids; ++s)
    {
        const idgx_metadata &m = md[s];
        const long long row0 = static_cast<long long>(m.baseline_offset) - md[0].baseline_offset + m.time_offset;
        const int c0 = ch ? ch[s].channel_begin : 0;
        const int nc = ch ? ch[s].nr_channels : plan->nr_channels;
        if (m.nr_timesteps < 0 || m.nr_timesteps > plan->max_nr_timesteps || row0 < 0 ||
            row0 + m.nr_timesteps > plan->nr_rows || c0 < 0 || nc < 0 || c0 + nc > plan->nr_channels ||
            m.station1 >= static_cast<uint32_t>(plan->nr_stations) ||
            m.station2 >= static_cast<uint32_t>(plan->nr_stations) || m.aterm_index < 0 ||
            m.aterm_index >= plan->nr_aterm_slots)
            return IDGX_ERR_INVALID_METADATA;
    }
    return IDGX_OK;
}

extern "C" int idgx_grid(const idgx_plan *plan, const idgx_metadata *d_metadata, const idgx_channels *d_channels,
                         const float *d_uvw, const float *d_wavenumbers, const void *d_visibilities,
                         const float *d_spheroidal, const void *d_aterms, void *d_subgrids, void *stream)
{
    const int rc = check_plan_struct(plan);
    if (rc != IDGX_OK)
        return rc;
    if (plan->nr_subgrids == 0)
        return IDGX_OK;
    if (!d_metadata || !d_uvw || !d_wavenumbers || !d_visibilities || !d_spheroidal || !d_aterms || !d_subgrids)
        return IDGX_ERR_INVALID_ARGUMENT;
    KernelArgs a{};
    a.p = to_dev(plan);
    a.metadata = d_metadata;
    a.channels = d_channels;
    a.uvw = d_uvw;
    a.wavenumbers = d_wavenumbers;
    a.visibilities = d_visibilities;
    a.spheroidal = d_spheroidal;
    a.aterms = d_aterms;
    a.subgrids_out = d_subgrids;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    switch (plan->precision)
    {
    case IDGX_PLAIN_FP32:
        return to_status(idgx::launch_grid_plain(a, accurate_sincos(plan), st));
    default:
        return to_status(idgx::launch_grid_mixed(a, plan->precision, (plan->flags & IDGX_FLAG_PHASE_MFMA) != 0, st));
    }
}

extern "C" int idgx_degrid(const idgx_plan *plan, const idgx_metadata *d_metadata, const idgx_channels *d_channels,
                           const float *d_uvw, const float *d_wavenumbers, const void *d_subgrids,
                           const float *d_spheroidal, const void *d_aterms, void *d_visibilities, void *stream)
{
    const int rc = check_plan_struct(plan);
    if (rc != IDGX_OK)
        return rc;
    if (plan->nr_subgrids == 0)
        return IDGX_OK;
    if (!d_metadata || !d_uvw || !d_wavenumbers || !d_visibilities || !d_spheroidal || !d_aterms || !d_subgrids)
        return IDGX_ERR_INVALID_ARGUMENT;
    KernelArgs a{};
    a.p = to_dev(plan);
    a.metadata = d_metadata;
    a.channels = d_channels;
    a.uvw = d_uvw;
    a.wavenumbers = d_wavenumbers;
    a.subgrids = d_subgrids;
    a.spheroidal = d_spheroidal;
    a.aterms = d_aterms;
    a.visibilities_out = d_visibilities;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    switch (plan->precision)
    {
    case IDGX_PLAIN_FP32:
        return to_status(idgx::launch_degrid_plain(a, accurate_sincos(plan), st));
    default:
        return to_status(idgx::launch_degrid_mixed(a, plan->precision, (plan->flags & IDGX_FLAG_PHASE_MFMA) != 0, st));
    }
}

extern "C" int idgx_subgrid_fft(const idgx_plan *plan, void *d_subgrids, int direction, void *stream)
{
    const int rc = check_plan_struct(plan);
    if (rc != IDGX_OK)
        return rc;
    if (direction != IDGX_FFT_FORWARD && direction != IDGX_FFT_BACKWARD)
        return IDGX_ERR_INVALID_ARGUMENT;
    if (!idgx::subgrid_fft_supported(plan->subgrid_size))
        return IDGX_ERR_UNSUPPORTED;
    if (plan->nr_subgrids == 0)
        return IDGX_OK;
    if (!d_subgrids)
        return IDGX_ERR_INVALID_ARGUMENT;
    return to_status(idgx::launch_subgrid_fft(d_subgrids, plan->nr_subgrids, plan->subgrid_size,
                                              direction == IDGX_FFT_FORWARD, static_cast<hipStream_t>(stream)));
}

namespace
{
// Tiles keyed by (colour, layer, tile row, tile column); sorting by key puts each colour's tiles together, and two
// tiles of one colour never touch the same cell of the same layer.
int build_adder_plan(const idgx_plan *plan, const idgx_metadata *md, int z_min, int nr_layers, idgx_adder_plan *ap,
                     int32_t *order, int32_t *tile_start)
{
    const int rc = check_plan_struct(plan);
    if (rc != IDGX_OK)
        return rc;
    if (!ap || nr_layers < 1 || (plan->nr_subgrids > 0 && (!md || !order || !tile_start)))
        return IDGX_ERR_INVALID_ARGUMENT;
    const int n = plan->nr_subgrids;
    const int T = plan->subgrid_size;
    auto floordiv = [](long long a, long long b) { return a >= 0 ? a / b : -((-a + b - 1) / b); };
    std::vector<std::pair<long long, int>> keys(n);
    const long long ntiles_axis = floordiv(plan->grid_size, T) + 3; // tiles from -1 .. G/T + 1
    const long long L = nr_layers;
    for (int s = 0; s < n; ++s)
    {
        const long long layer = nr_layers > 1 ? static_cast<long long>(md[s].z) - z_min : 0;
        if (layer < 0 || layer >= L)
            return IDGX_ERR_INVALID_METADATA;
        const long long tx = floordiv(md[s].x, T) + 1, ty = floordiv(md[s].y, T) + 1;
        const long long color = ((tx - 1) & 1) + 2 * ((ty - 1) & 1);
        const long long tx_c = std::min(std::max(tx, 0LL), ntiles_axis - 1);
        const long long ty_c = std::min(std::max(ty, 0LL), ntiles_axis - 1);
        keys[s] = {((color * L + layer) * ntiles_axis + ty_c) * ntiles_axis + tx_c, s};
    }
    std::stable_sort(keys.begin(), keys.end(),
                     [](const std::pair<long long, int> &a, const std::pair<long long, int> &b) {
                         return a.first < b.first;
                     });
    ap->tile_size = T;
    ap->z_min = nr_layers > 1 ? z_min : 0;
    ap->nr_layers = nr_layers;
    // Tiles (runs of equal keys): colour, first entry, count.
    struct Tile
    {
        int color, first, count;
    };
    std::vector<Tile> runs;
    for (int i = 0; i < n; ++i)
    {
        if (i == 0 || keys[i].first != keys[i - 1].first)
            runs.push_back({static_cast<int>(keys[i].first / (L * ntiles_axis * ntiles_axis)), i, 0});
        ++runs.back().count;
    }
    // Within a colour, heaviest tile first: a tile's workgroups walk its subgrids one after another, so the crowded
    // tiles near the uv centre start at the head of the launch instead of finishing it.  The order of the subgrids
    // inside a tile, and so every cell's summation order, is unchanged.
    std::stable_sort(runs.begin(), runs.end(), [](const Tile &a, const Tile &b) {
        return a.color != b.color ? a.color < b.color : a.count > b.count;
    });
    for (int c = 0; c <= 4; ++c)
        ap->color_begin[c] = 0;
    int tiles = 0, pos = 0;
    for (const Tile &t : runs)
    {
        tile_start[tiles] = pos;
        for (int k = 0; k < t.count; ++k)
            order[pos++] = keys[t.first + k].second;
        for (int c = t.color + 1; c <= 4; ++c)
            ap->color_begin[c] = tiles + 1;
        ++tiles;
    }
    tile_start[tiles] = n;
    ap->nr_tiles = tiles;
    for (int c = 1; c <= 4; ++c)
        ap->color_begin[c] = std::max(ap->color_begin[c], ap->color_begin[c - 1]);
    ap->color_begin[4] = tiles;
    return IDGX_OK;
}
} // namespace

extern "C" int idgx_adder_plan_build(const idgx_plan *plan, const idgx_metadata *md, idgx_adder_plan *ap,
                                     int32_t *order, int32_t *tile_start)
{
    return build_adder_plan(plan, md, 0, 1, ap, order, tile_start);
}

extern "C" int idgx_adder_plan_build_layered(const idgx_plan *plan, const idgx_metadata *md, int z_min,
                                             int nr_layers, idgx_adder_plan *ap, int32_t *order, int32_t *tile_start)
{
    return build_adder_plan(plan, md, z_min, nr_layers, ap, order, tile_start);
}

extern "C" int idgx_add_subgrids(const idgx_plan *plan, const idgx_adder_plan *ap, const idgx_metadata *d_metadata,
                                 const int32_t *d_order, const int32_t *d_tile_start, const void *d_subgrids,
                                 void *d_grid, void *stream)
{
    const int rc = check_plan_struct(plan);
    if (rc != IDGX_OK)
        return rc;
    if (!ap || ap->tile_size < plan->subgrid_size || ap->nr_tiles < 0 || ap->color_begin[0] != 0 ||
        ap->color_begin[4] != ap->nr_tiles || ap->nr_layers < 1)
        return IDGX_ERR_INVALID_ARGUMENT;
    for (int c = 0; c < 4; ++c)
        if (ap->color_begin[c] > ap->color_begin[c + 1])
            return IDGX_ERR_INVALID_ARGUMENT;
    if (plan->nr_subgrids == 0 || ap->nr_tiles == 0)
        return IDGX_OK;
    if (!d_metadata || !d_order || !d_tile_start || !d_subgrids || !d_grid)
        return IDGX_ERR_INVALID_ARGUMENT;
    return to_status(idgx::launch_adder(d_metadata, d_order, d_tile_start, ap->color_begin, plan->subgrid_size,
                                        ap->tile_size, plan->grid_size, ap->z_min, ap->nr_layers, d_subgrids, d_grid,
                                        static_cast<hipStream_t>(stream)));
}

extern "C" int idgx_split_subgrids(const idgx_plan *plan, const idgx_metadata *d_metadata, const void *d_grid,
                                   void *d_subgrids, void *stream)
{
    const int rc = check_plan_struct(plan);
    if (rc != IDGX_OK)
        return rc;
    if (plan->nr_subgrids == 0)
        return IDGX_OK;
    if (!d_metadata || !d_grid || !d_subgrids)
        return IDGX_ERR_INVALID_ARGUMENT;
    return to_status(idgx::launch_splitter(d_metadata, plan->nr_subgrids, plan->subgrid_size, plan->grid_size, 0, 1,
                                           d_grid, d_subgrids, static_cast<hipStream_t>(stream)));
}

extern "C" int idgx_split_subgrids_layered(const idgx_plan *plan, const idgx_metadata *d_metadata, int z_min,
                                           int nr_layers, const void *d_grid, void *d_subgrids, void *stream)
{
    const int rc = check_plan_struct(plan);
    if (rc != IDGX_OK)
        return rc;
    if (nr_layers < 1)
        return IDGX_ERR_INVALID_ARGUMENT;
    if (plan->nr_subgrids == 0)
        return IDGX_OK;
    if (!d_metadata || !d_grid || !d_subgrids)
        return IDGX_ERR_INVALID_ARGUMENT;
    return to_status(idgx::launch_splitter(d_metadata, plan->nr_subgrids, plan->subgrid_size, plan->grid_size,
                                           nr_layers > 1 ? z_min : 0, nr_layers, d_grid, d_subgrids,
                                           static_cast<hipStream_t>(stream)));
}

extern "C" int idgx_grid_fft(const idgx_plan *plan, void *d_grid, int nr_layers, int direction, void *stream)
{
    const int rc = check_plan_struct(plan);
    if (rc != IDGX_OK)
        return rc;
    if (nr_layers < 1 || !d_grid || (direction != IDGX_FFT_FORWARD && direction != IDGX_FFT_BACKWARD))
        return IDGX_ERR_INVALID_ARGUMENT;
    if (!idgx::grid_fft_supported(plan->grid_size))
        return IDGX_ERR_UNSUPPORTED;
    return to_status(idgx::launch_grid_fft(d_grid, plan->grid_size, nr_layers, direction == IDGX_FFT_FORWARD,
                                           static_cast<hipStream_t>(stream)));
}

extern "C" int idgx_grid_fft_noalloc(const idgx_plan *plan, void *d_grid, int nr_layers, int direction, void *stream)
{
    const int rc = check_plan_struct(plan);
    if (rc != IDGX_OK)
        return rc;
    if (nr_layers < 1 || !d_grid || (direction != IDGX_FFT_FORWARD && direction != IDGX_FFT_BACKWARD))
        return IDGX_ERR_INVALID_ARGUMENT;
    if (!idgx::grid_fft_noalloc_supported(plan->grid_size))
        return IDGX_ERR_UNSUPPORTED; // G outside 1024, 2048, 4096: the scratch would not fit the fixed array
    return to_status(idgx::launch_grid_fft_noalloc(d_grid, plan->grid_size, nr_layers, direction == IDGX_FFT_FORWARD,
                                                   static_cast<hipStream_t>(stream)));
}

extern "C" int idgx_grid_box(const idgx_plan *plan, const idgx_metadata *h_metadata, int32_t box[4])
{
    const int rc = check_plan_struct(plan);
    if (rc != IDGX_OK)
        return rc;
    if (!box || (plan->nr_subgrids > 0 && !h_metadata))
        return IDGX_ERR_INVALID_ARGUMENT;
    const long long G = plan->grid_size, S = plan->subgrid_size;
    box[0] = box[1] = box[2] = box[3] = 0;
    if (plan->nr_subgrids == 0)
        return IDGX_OK;
    long long x0 = G, x1 = 0, y0 = G, y1 = 0;
    for (int s = 0; s < plan->nr_subgrids; ++s)
    {
        const long long x = h_metadata[s].x, y = h_metadata[s].y;
        x0 = std::min(x0, x);
        x1 = std::max(x1, x + S);
        y0 = std::min(y0, y);
        y1 = std::max(y1, y + S);
    }
    // Clamped to the grid (the adder and splitter skip the cells of a subgrid that overhang it).
    x0 = std::max(0LL, std::min(x0, G));
    y0 = std::max(0LL, std::min(y0, G));
    x1 = std::max(x0, std::min(x1, G));
    y1 = std::max(y0, std::min(y1, G));
    box[0] = static_cast<int32_t>(x0);
    box[1] = static_cast<int32_t>(x1);
    box[2] = static_cast<int32_t>(y0);
    box[3] = static_cast<int32_t>(y1);
    return IDGX_OK;
}

extern "C" int idgx_grid_fft_box(const idgx_plan *plan, void *d_grid, int nr_layers, int direction,
                                 const int32_t box[4], void *stream)
{
    const int rc = check_plan_struct(plan);
    if (rc != IDGX_OK)
        return rc;
    const int G = plan->grid_size;
    if (nr_layers < 1 || !d_grid || !box || (direction != IDGX_FFT_FORWARD && direction != IDGX_FFT_BACKWARD) ||
        box[0] < 0 || box[0] > box[1] || box[1] > G || box[2] < 0 || box[2] > box[3] || box[3] > G)
        return IDGX_ERR_INVALID_ARGUMENT;
    if (!idgx::grid_fft_supported(G))
        return IDGX_ERR_UNSUPPORTED;
    const int b[4] = {box[0], box[1], box[2], box[3]};
    return to_status(idgx::launch_grid_fft(d_grid, G, nr_layers, direction == IDGX_FFT_FORWARD,
                                           static_cast<hipStream_t>(stream), b));
}

extern "C" int idgx_grid_clear_box(const idgx_plan *plan, void *d_grid, int nr_layers, const int32_t box[4],
                                   void *stream)
{
    const int rc = check_plan_struct(plan);
    if (rc != IDGX_OK)
        return rc;
    const int G = plan->grid_size;
    if (nr_layers < 1 || !d_grid || !box || box[0] < 0 || box[0] > box[1] || box[1] > G || box[2] < 0 ||
        box[2] > box[3] || box[3] > G)
        return IDGX_ERR_INVALID_ARGUMENT;
    if (!idgx::grid_fft_supported(G))
        return IDGX_ERR_UNSUPPORTED;
    const int b[4] = {box[0], box[1], box[2], box[3]};
    return to_status(idgx::launch_grid_clear_box(d_grid, G, nr_layers, b, static_cast<hipStream_t>(stream)));
}

extern "C" int idgx_fft2d(const void *d_in, void *d_out, int nx, int ny, int64_t batch, int sign, int precision,
                          void *stream)
{
    if (nx < 1 || ny < 1 || batch < 1 || (sign != 1 && sign != -1) ||
        (precision != IDGX_FFT_FP32 && precision != IDGX_FFT_FP64))
        return IDGX_ERR_INVALID_ARGUMENT;
    if (!idgx::fft2d_supported(nx, ny) || batch > 0x7fffffffLL)
        return IDGX_ERR_UNSUPPORTED;
    if (!d_in || !d_out)
        return IDGX_ERR_INVALID_ARGUMENT;
    return to_status(idgx::launch_fft2d(d_in, d_out, nx, ny, batch, sign, precision == IDGX_FFT_FP64,
                                        static_cast<hipStream_t>(stream)));
}

extern "C" int idgx_fft2d_own_kernels(int nx, int ny, int precision)
{
    if (nx < 1 || ny < 1 || (precision != IDGX_FFT_FP32 && precision != IDGX_FFT_FP64))
        return -IDGX_ERR_INVALID_ARGUMENT;
    return idgx::fft2d_own_kernels(nx, ny, precision == IDGX_FFT_FP64) ? 1 : 0;
}

extern "C" int idgx_wstack_to_image(const idgx_plan *plan, const void *d_layers, int z_min, int nr_layers,
                                    void *d_image, void *stream)
{
    const int rc = check_plan_struct(plan);
    if (rc != IDGX_OK)
        return rc;
    if (nr_layers < 1 || !d_layers || !d_image || d_layers == d_image)
        return IDGX_ERR_INVALID_ARGUMENT;
    return to_status(idgx::launch_wstack(d_layers, d_image, plan->grid_size, z_min, nr_layers, plan->image_size,
                                         plan->w_step_in_lambda, true, static_cast<hipStream_t>(stream)));
}

extern "C" int idgx_wstack_from_image(const idgx_plan *plan, const void *d_image, int z_min, int nr_layers,
                                      void *d_layers, void *stream)
{
    const int rc = check_plan_struct(plan);
    if (rc != IDGX_OK)
        return rc;
    if (nr_layers < 1 || !d_layers || !d_image || d_layers == d_image)
        return IDGX_ERR_INVALID_ARGUMENT;
    return to_status(idgx::launch_wstack(d_image, d_layers, plan->grid_size, z_min, nr_layers, plan->image_size,
                                         plan->w_step_in_lambda, false, static_cast<hipStream_t>(stream)));
}

namespace
{
int grid_wstack(const idgx_plan *plan, void *d_grid, void *d_image, int z_min, int nr_layers, const int32_t *box,
                bool to_image, void *stream)
{
    const int rc = check_plan_struct(plan);
    if (rc != IDGX_OK)
        return rc;
    const int G = plan->grid_size;
    if (nr_layers < 1 || !d_grid || !d_image || d_grid == d_image || !box || box[0] < 0 || box[0] > box[1] ||
        box[1] > G || box[2] < 0 || box[2] > box[3] || box[3] > G)
        return IDGX_ERR_INVALID_ARGUMENT;
    if (!idgx::grid_fft_supported(G))
        return IDGX_ERR_UNSUPPORTED;
    const int b[4] = {box[0], box[1], box[2], box[3]};
    return to_status(idgx::launch_grid_wstack(d_grid, d_image, G, z_min, nr_layers, plan->image_size,
                                              plan->w_step_in_lambda, to_image, b, static_cast<hipStream_t>(stream)));
}
} // namespace

extern "C" int idgx_grid_to_image(const idgx_plan *plan, void *d_grid, int z_min, int nr_layers, const int32_t box[4],
                                  void *d_image, void *stream)
{
    return grid_wstack(plan, d_grid, d_image, z_min, nr_layers, box, true, stream);
}

extern "C" int idgx_image_to_grid(const idgx_plan *plan, const void *d_image, int z_min, int nr_layers,
                                  const int32_t box[4], void *d_grid, void *stream)
{
    return grid_wstack(plan, d_grid, const_cast<void *>(d_image), z_min, nr_layers, box, false, stream);
}

extern "C" size_t idgx_plan_workspace_bytes(int64_t nr_segments)
{
    return nr_segments < 0 ? 0 : idgx::plan_workspace_bytes(nr_segments);
}

extern "C" int idgx_plan_segments(const idgx_planner *pl, const float *d_uvw, const double *d_frequencies,
                                  const uint32_t *d_station1, const uint32_t *d_station2,
                                  const int32_t *d_seg_baseline, const int32_t *d_seg_t0, const int32_t *d_seg_t1,
                                  int64_t nr_segments, void *d_workspace, idgx_metadata *d_metadata,
                                  idgx_channels *d_channels, int64_t capacity, int64_t *h_nr_subgrids, void *stream)
{
    if (!pl || !h_nr_subgrids || nr_segments < 0 || capacity < 0 || pl->subgrid_size < 2 || pl->grid_size < 1 ||
        pl->nr_times < 1 || pl->nr_channels < 1 || pl->aterm_interval < 0 || pl->support < 0 ||
        pl->support >= pl->subgrid_size || !(pl->image_size > 0) || !(pl->w_step_in_lambda >= 0))
        return IDGX_ERR_INVALID_ARGUMENT;
    if (nr_segments > 0 && (!d_uvw || !d_frequencies || !d_station1 || !d_station2 || !d_seg_baseline || !d_seg_t0 ||
                            !d_seg_t1 || !d_workspace || (capacity > 0 && (!d_metadata || !d_channels))))
        return IDGX_ERR_INVALID_ARGUMENT;
    long long count = 0;
    if (nr_segments == 0)
    {
        *h_nr_subgrids = 0;
        return IDGX_OK;
    }
    const int rc = to_status(idgx::launch_plan(*pl, d_uvw, d_frequencies, d_station1, d_station2, d_seg_baseline,
                                               d_seg_t0, d_seg_t1, nr_segments, d_workspace, d_metadata, d_channels,
                                               capacity, &count, static_cast<hipStream_t>(stream)));
    *h_nr_subgrids = count;
    return rc;
}

extern "C" int idgx_plain_kernel_shape(const idgx_plan *plan, int has_channel_windows, int *gridder_pixels_per_thread,
                                       int *degridder_channels_per_thread)
{
    const int rc = check_plan_struct(plan);
    if (rc != IDGX_OK)
        return rc;
    if (gridder_pixels_per_thread)
        *gridder_pixels_per_thread = idgx::plain_grid_pixels_for(plan->subgrid_size);
    if (degridder_channels_per_thread)
        *degridder_channels_per_thread = idgx::plain_degrid_channels_for(has_channel_windows != 0, plan->nr_channels);
    return IDGX_OK;
}

extern "C" const char *idgx_strerror(int status)
{
    switch (status)
    {
    case IDGX_OK:
        return "ok";
    case IDGX_ERR_INVALID_ARGUMENT:
        return "invalid argument";
    case IDGX_ERR_INVALID_METADATA:
        return "subgrid metadata reads outside the arrays";
    case IDGX_ERR_UNSUPPORTED:
        return "unsupported request";
    case IDGX_ERR_DEVICE:
        return "HIP runtime error";
    case IDGX_ERR_OUT_OF_MEMORY:
        return "out of device memory";
    default:
        return "unknown status";
    }
}

extern "C" const char *idgx_version(void)
{
    return "idgx 0.1 gfx950";
}
