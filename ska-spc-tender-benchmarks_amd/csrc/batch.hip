// Batched execution of a whole plan: every subgrid of up to four subgrid classes (one subgrid size each) runs in
// one persistent gridder (or degridder) launch fed from a device work queue, optionally captured in a hipGraph,
// with the tile shape chosen by an autotuner that times the compiled shapes on the real data.
//
// Work queue.  An item is (class, subgrid, block): a block of pixels for the gridder, a chunk of visibilities
// for the degridder.  Items are built on the host from the plan (metadata, channel windows and visibility
// selections give every subgrid's visibility count V_s), costed as visibilities x pixels, and sorted heaviest
// first (longest-processing-time order), so the uneven tail (flagging, ragged timesteps, mixed sizes) finishes
// together.  Persistent workgroups (as many as fit on the chip) take items through one atomic counter, which a
// one-thread kernel clears before every launch (no memset nodes, so the sequence can be captured).
//
// Degridder pixels: before the persistent degridder, one workgroup per subgrid applies taper and A-terms and
// splits the pixels into the packed words the degridder stages (32 B per pixel slot), so the work items copy
// them instead of recomputing them for every chunk of visibilities.
//
// Device memory (queues, counters, the prepared pixels: 32 B per pixel per subgrid) is allocated once in
// idgx_batch_create; the execution entry points allocate nothing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "idgx.h"
#include "kernels.hpp"

struct idgx_batch
{
    int nr_classes = 0;
    idgx_plan plans[idgx::kMaxClasses];
    std::vector<int64_t> vis_count[idgx::kMaxClasses]; // V_s per subgrid
    int precision = IDGX_MIXED_BF16X3;
    int grid_shape = 0, degrid_shape = 0;
    int grid_wg = 0, degrid_wg = 0;
    int nr_grid_items = 0, nr_degrid_items = 0;
    size_t grid_capacity = 0, degrid_capacity = 0;
    uint2 *d_grid_items = nullptr;
    uint2 *d_degrid_items = nullptr;
    int *d_counters = nullptr; // [0] gridder, [1] degridder
    // Degridder pixels prepared once per subgrid and pass (per class): split words, FP16 scales, slot n table.
    uint4 *d_prepared[idgx::kMaxClasses] = {};
    int *d_prepared_scale[idgx::kMaxClasses] = {};
    float *d_slot_n[idgx::kMaxClasses] = {};
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
};

namespace
{
using idgx::KernelArgs;
using idgx::QueueArgs;

int to_status(hipError_t e)
{
    if (e == hipSuccess)
        return IDGX_OK;
    if (e == hipErrorOutOfMemory)
        return IDGX_ERR_OUT_OF_MEMORY;
    return IDGX_ERR_DEVICE;
}

void drop_graph(idgx_batch *b)
{
    if (b->graph_exec)
        (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph)
        (void)hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
}

struct Item
{
    uint2 v;
    double cost;
};

std::vector<Item> build_items(const idgx_batch *b, bool gridder, int shape)
{
    std::vector<Item> items;
    const long long per = idgx::degrid_visibilities_per_item(shape);
    for (int k = 0; k < b->nr_classes; ++k)
    {
        const int S = b->plans[k].subgrid_size;
        const long long P = static_cast<long long>(S) * S;
        const int gitems = idgx::grid_items_per_subgrid(shape, S);
        std::vector<double> gpix(gitems);
        for (int i = 0; i < gitems; ++i)
            gpix[i] = static_cast<double>(idgx::grid_item_pixels(shape, S, i));
        for (int s = 0; s < b->plans[k].nr_subgrids; ++s)
        {
            const long long V = b->vis_count[k][s];
            const uint32_t tag = (static_cast<uint32_t>(k) << 28) | static_cast<uint32_t>(s);
            if (gridder)
            {
                // Every pixel block is written (zeros for an empty subgrid); the epilogue costs ~ V of 1.
                for (int blk = 0; blk < gitems; ++blk)
                    items.push_back({{tag, static_cast<uint32_t>(blk)}, gpix[blk] * static_cast<double>(V + 1)});
            }
            else
            {
                for (long long v0 = 0, blk = 0; v0 < V; v0 += per, ++blk)
                    items.push_back({{tag, static_cast<uint32_t>(blk)},
                                     static_cast<double>(std::min(per, V - v0)) * static_cast<double>(P)});
            }
        }
    }
    std::stable_sort(items.begin(), items.end(), [](const Item &x, const Item &y) { return x.cost > y.cost; });
    return items;
}

size_t max_items(const idgx_batch *b, bool gridder)
{
    size_t m = 0;
    for (int sh = 0; sh < idgx::mixed_shape_count(); ++sh)
        m = std::max(m, build_items(b, gridder, sh).size());
    return m;
}

int upload_queue(idgx_batch *b, bool gridder, int shape)
{
    const std::vector<Item> items = build_items(b, gridder, shape);
    std::vector<uint2> flat(items.size());
    for (size_t i = 0; i < items.size(); ++i)
        flat[i] = items[i].v;
    uint2 *dst = gridder ? b->d_grid_items : b->d_degrid_items;
    if (flat.size() > (gridder ? b->grid_capacity : b->degrid_capacity))
        return IDGX_ERR_INVALID_ARGUMENT;
    if (!flat.empty())
    {
        const hipError_t e = hipMemcpy(dst, flat.data(), flat.size() * sizeof(uint2), hipMemcpyHostToDevice);
        if (e != hipSuccess)
            return to_status(e);
    }
    (gridder ? b->nr_grid_items : b->nr_degrid_items) = static_cast<int>(flat.size());
    (gridder ? b->grid_shape : b->degrid_shape) = shape;
    (gridder ? b->grid_wg : b->degrid_wg) = idgx::queue_workgroups(b->precision, gridder, shape);
    return IDGX_OK;
}

int make_queue_args(const idgx_batch *b, const idgx_batch_data *d, bool gridder, QueueArgs &q)
{
    if (!d || !d->d_uvw || !d->d_wavenumbers || (gridder ? !d->d_visibilities : !d->d_visibilities_out))
        return IDGX_ERR_INVALID_ARGUMENT;
    std::memset(&q, 0, sizeof(q));
    for (int k = 0; k < b->nr_classes; ++k)
    {
        const idgx_plan &p = b->plans[k];
        const idgx_batch_class_data &c = d->cls[k];
        if (p.nr_subgrids > 0 &&
            (!c.d_metadata || !c.d_spheroidal || !c.d_aterms || !c.d_subgrids || (!c.d_vis_index != !c.d_vis_offset)))
            return IDGX_ERR_INVALID_ARGUMENT;
        KernelArgs &a = q.cls[k];
        a.p.nr_subgrids = p.nr_subgrids;
        a.p.S = p.subgrid_size;
        a.p.G = p.grid_size;
        a.p.nr_channels = p.nr_channels;
        a.p.nr_stations = p.nr_stations;
        a.p.nr_aterm_slots = p.nr_aterm_slots;
        a.p.max_nr_timesteps = p.max_nr_timesteps;
        a.p.nr_rows = p.nr_rows;
        a.p.image_size = p.image_size;
        a.p.w_step = p.w_step_in_lambda;
        a.p.convention = p.aterm_convention;
        a.metadata = c.d_metadata;
        a.channels = c.d_channels;
        a.uvw = d->d_uvw;
        a.wavenumbers = d->d_wavenumbers;
        a.visibilities = d->d_visibilities;
        a.visibilities_out = d->d_visibilities_out;
        a.spheroidal = c.d_spheroidal;
        a.aterms = c.d_aterms;
        a.subgrids = c.d_subgrids;
        a.subgrids_out = c.d_subgrids;
        a.vis_index = c.d_vis_index;
        a.vis_offset = reinterpret_cast<const long long *>(c.d_vis_offset);
        if (!gridder)
        {
            a.prepared = b->d_prepared[k];
            a.prepared_scale = b->d_prepared_scale[k];
            a.slot_n = b->d_slot_n[k];
        }
    }
    q.items = gridder ? b->d_grid_items : b->d_degrid_items;
    q.nr_items = gridder ? b->nr_grid_items : b->nr_degrid_items;
    q.counter = b->d_counters + (gridder ? 0 : 1);
    return IDGX_OK;
}

int run(idgx_batch *b, const idgx_batch_data *d, bool gridder, hipStream_t st)
{
    QueueArgs q;
    int rc = make_queue_args(b, d, gridder, q);
    if (rc != IDGX_OK)
        return rc;
    if (q.nr_items == 0)
        return IDGX_OK;
    hipError_t e = hipSuccess;
    for (int k = 0; k < b->nr_classes && !gridder && e == hipSuccess; ++k)
        e = idgx::launch_degrid_prepare(q.cls[k], b->precision, st);
    if (e == hipSuccess)
        e = idgx::launch_reset_counter(q.counter, st);
    if (e != hipSuccess)
        return to_status(e);
    e = gridder ? idgx::launch_grid_mixed_queue(q, b->precision, b->grid_shape, b->grid_wg, st)
                : idgx::launch_degrid_mixed_queue(q, b->precision, b->degrid_shape, b->degrid_wg, st);
    return to_status(e);
}
} // namespace

extern "C" int idgx_batch_create(const idgx_batch_class *classes, int nr_classes, idgx_batch **out)
{
    if (!out)
        return IDGX_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!classes || nr_classes < 1 || nr_classes > idgx::kMaxClasses)
        return IDGX_ERR_INVALID_ARGUMENT;
    const idgx_plan &p0 = classes[0].plan;
    for (int k = 0; k < nr_classes; ++k)
    {
        const idgx_batch_class &c = classes[k];
        const int rc = idgx_check_plan(&c.plan, c.h_metadata, c.h_channels);
        if (rc != IDGX_OK)
            return rc;
        if (c.plan.precision != p0.precision || c.plan.nr_channels != p0.nr_channels || c.plan.nr_rows != p0.nr_rows ||
            c.plan.nr_stations != p0.nr_stations || c.plan.grid_size != p0.grid_size ||
            c.plan.image_size != p0.image_size)
            return IDGX_ERR_INVALID_ARGUMENT; // classes differ only in subgrid size / subgrids / A-term slots
        if (c.plan.nr_subgrids >= (1 << 28))
            return IDGX_ERR_UNSUPPORTED;
    }
    if (p0.precision == IDGX_PLAIN_FP32)
        return IDGX_ERR_UNSUPPORTED; // the persistent forms are the mixed-precision kernels
    idgx_batch *b = new (std::nothrow) idgx_batch();
    if (!b)
        return IDGX_ERR_OUT_OF_MEMORY;
    b->nr_classes = nr_classes;
    b->precision = p0.precision;
    for (int k = 0; k < nr_classes; ++k)
    {
        const idgx_batch_class &c = classes[k];
        b->plans[k] = c.plan;
        b->vis_count[k].resize(c.plan.nr_subgrids);
        for (int s = 0; s < c.plan.nr_subgrids; ++s)
        {
            const int nc = c.h_channels ? c.h_channels[s].nr_channels : c.plan.nr_channels;
            long long v = static_cast<long long>(c.h_metadata[s].nr_timesteps) * nc;
            if (c.h_vis_offset)
            {
                v = c.h_vis_offset[s + 1] - c.h_vis_offset[s];
                if (v < 0 || c.h_vis_offset[s] < 0 || c.h_vis_offset[s + 1] > c.nr_vis_index ||
                    v > static_cast<long long>(c.h_metadata[s].nr_timesteps) * nc)
                {
                    delete b;
                    return IDGX_ERR_INVALID_METADATA;
                }
            }
            b->vis_count[k][s] = v;
        }
    }
    b->grid_capacity = std::max<size_t>(1, max_items(b, true));
    b->degrid_capacity = std::max<size_t>(1, max_items(b, false));
    hipError_t e = hipMalloc(&b->d_grid_items, b->grid_capacity * sizeof(uint2));
    if (e == hipSuccess)
        e = hipMalloc(&b->d_degrid_items, b->degrid_capacity * sizeof(uint2));
    if (e == hipSuccess)
        e = hipMalloc(&b->d_counters, 2 * sizeof(int));
    for (int k = 0; k < nr_classes && e == hipSuccess; ++k)
    {
        const int S = b->plans[k].subgrid_size;
        const long long n = std::max(1, b->plans[k].nr_subgrids);
        e = hipMalloc(&b->d_prepared[k], idgx::degrid_prepared_bytes(S, n));
        if (e == hipSuccess)
            e = hipMalloc(&b->d_prepared_scale[k], n * sizeof(int));
        if (e == hipSuccess)
            e = hipMalloc(&b->d_slot_n[k], static_cast<size_t>(S) * idgx::degrid_row_stride(S) * sizeof(float));
    }
    int rc = to_status(e);
    if (rc == IDGX_OK)
        rc = upload_queue(b, true, idgx::mixed_default_shape(true));
    if (rc == IDGX_OK)
        rc = upload_queue(b, false, idgx::mixed_default_shape(false));
    if (rc != IDGX_OK)
    {
        idgx_batch_destroy(b);
        return rc;
    }
    *out = b;
    return IDGX_OK;
}

extern "C" int idgx_batch_shape_count(void)
{
    return idgx::mixed_shape_count();
}

extern "C" int idgx_batch_shape_info(int shape, int *tiles_per_wave, int *lds_batch, int *phase_mfma)
{
    if (shape < 0 || shape >= idgx::mixed_shape_count())
        return IDGX_ERR_INVALID_ARGUMENT;
    const idgx::MixedShape s = idgx::mixed_shape(shape);
    if (tiles_per_wave)
        *tiles_per_wave = s.tiles;
    if (lds_batch)
        *lds_batch = s.batch;
    if (phase_mfma)
        *phase_mfma = s.phase_mfma;
    return IDGX_OK;
}

extern "C" int idgx_batch_set_shape(idgx_batch *b, int gridder_shape, int degridder_shape)
{
    if (!b || gridder_shape < 0 || gridder_shape >= idgx::mixed_shape_count() || degridder_shape < 0 ||
        degridder_shape >= idgx::mixed_shape_count())
        return IDGX_ERR_INVALID_ARGUMENT;
    drop_graph(b);
    int rc = IDGX_OK;
    if (gridder_shape != b->grid_shape)
        rc = upload_queue(b, true, gridder_shape);
    if (rc == IDGX_OK && degridder_shape != b->degrid_shape)
        rc = upload_queue(b, false, degridder_shape);
    return rc;
}

extern "C" int idgx_batch_get_shape(const idgx_batch *b, int *gridder_shape, int *degridder_shape)
{
    if (!b)
        return IDGX_ERR_INVALID_ARGUMENT;
    if (gridder_shape)
        *gridder_shape = b->grid_shape;
    if (degridder_shape)
        *degridder_shape = b->degrid_shape;
    return IDGX_OK;
}

extern "C" int idgx_batch_grid(idgx_batch *b, const idgx_batch_data *data, void *stream)
{
    if (!b)
        return IDGX_ERR_INVALID_ARGUMENT;
    return run(b, data, true, static_cast<hipStream_t>(stream));
}

extern "C" int idgx_batch_degrid(idgx_batch *b, const idgx_batch_data *data, void *stream)
{
    if (!b)
        return IDGX_ERR_INVALID_ARGUMENT;
    return run(b, data, false, static_cast<hipStream_t>(stream));
}

extern "C" int idgx_batch_autotune(idgx_batch *b, const idgx_batch_data *data, int repeats, void *stream,
                                   float *best_grid_ms, float *best_degrid_ms)
{
    if (!b || repeats < 1)
        return IDGX_ERR_INVALID_ARGUMENT;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    drop_graph(b);
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess)
        return IDGX_ERR_DEVICE;
    if (hipEventCreate(&e1) != hipSuccess)
    {
        (void)hipEventDestroy(e0);
        return IDGX_ERR_DEVICE;
    }
    int rc = IDGX_OK;
    for (int dir = 0; dir < 2 && rc == IDGX_OK; ++dir)
    {
        const bool gridder = dir == 0;
        float best = -1.F;
        int best_shape = 0;
        for (int sh = 0; sh < idgx::mixed_shape_count() && rc == IDGX_OK; ++sh)
        {
            rc = upload_queue(b, gridder, sh);
            if (rc != IDGX_OK)
                break;
            rc = run(b, data, gridder, st); // warm-up
            if (rc != IDGX_OK)
                break;
            if (hipEventRecord(e0, st) != hipSuccess)
                rc = IDGX_ERR_DEVICE;
            for (int r = 0; r < repeats && rc == IDGX_OK; ++r)
                rc = run(b, data, gridder, st);
            float ms = 0.F;
            if (rc == IDGX_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                                  hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                rc = IDGX_ERR_DEVICE;
            ms /= repeats;
            if (rc == IDGX_OK && (best < 0.F || ms < best))
            {
                best = ms;
                best_shape = sh;
            }
        }
        if (rc == IDGX_OK)
            rc = upload_queue(b, gridder, best_shape);
        if (gridder && best_grid_ms)
            *best_grid_ms = best;
        if (!gridder && best_degrid_ms)
            *best_degrid_ms = best;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

extern "C" int idgx_batch_capture(idgx_batch *b, const idgx_batch_data *data, int stages, void *stream)
{
    if (!b || !stream || stages < 1 || stages > (IDGX_STAGE_GRID | IDGX_STAGE_DEGRID))
        return IDGX_ERR_INVALID_ARGUMENT;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    drop_graph(b);
    QueueArgs probe;
    if ((stages & IDGX_STAGE_GRID) && make_queue_args(b, data, true, probe) != IDGX_OK)
        return IDGX_ERR_INVALID_ARGUMENT;
    if ((stages & IDGX_STAGE_DEGRID) && make_queue_args(b, data, false, probe) != IDGX_OK)
        return IDGX_ERR_INVALID_ARGUMENT;
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess)
        return to_status(e);
    int rc = IDGX_OK;
    if (stages & IDGX_STAGE_GRID)
        rc = run(b, data, true, st);
    if (rc == IDGX_OK && (stages & IDGX_STAGE_DEGRID))
        rc = run(b, data, false, st);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(st, &g);
    if (rc != IDGX_OK || e != hipSuccess)
    {
        if (g)
            (void)hipGraphDestroy(g);
        return rc != IDGX_OK ? rc : to_status(e);
    }
    b->graph = g;
    e = hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess)
    {
        drop_graph(b);
        return to_status(e);
    }
    return IDGX_OK;
}

extern "C" int idgx_batch_launch(idgx_batch *b, void *stream)
{
    if (!b || !b->graph_exec)
        return IDGX_ERR_INVALID_ARGUMENT;
    return to_status(hipGraphLaunch(b->graph_exec, static_cast<hipStream_t>(stream)));
}

extern "C" int idgx_batch_items(const idgx_batch *b, int *gridder_items, int *degridder_items, int *gridder_workgroups,
                                int *degridder_workgroups)
{
    if (!b)
        return IDGX_ERR_INVALID_ARGUMENT;
    if (gridder_items)
        *gridder_items = b->nr_grid_items;
    if (degridder_items)
        *degridder_items = b->nr_degrid_items;
    if (gridder_workgroups)
        *gridder_workgroups = b->grid_wg;
    if (degridder_workgroups)
        *degridder_workgroups = b->degrid_wg;
    return IDGX_OK;
}

extern "C" void idgx_batch_destroy(idgx_batch *b)
{
    if (!b)
        return;
    drop_graph(b);
    if (b->d_grid_items)
        (void)hipFree(b->d_grid_items);
    if (b->d_degrid_items)
        (void)hipFree(b->d_degrid_items);
    if (b->d_counters)
        (void)hipFree(b->d_counters);
    for (int k = 0; k < idgx::kMaxClasses; ++k)
    {
        if (b->d_prepared[k])
            (void)hipFree(b->d_prepared[k]);
        if (b->d_prepared_scale[k])
            (void)hipFree(b->d_prepared_scale[k]);
        if (b->d_slot_n[k])
            (void)hipFree(b->d_slot_n[k]);
    }
    delete b;
}
