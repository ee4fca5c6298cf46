// Batched execution of a whole plan: the subgrids of up to four subgrid classes (one subgrid size each) run in
// persistent gridder (or degridder) launches fed from device work queues, optionally captured in a hipGraph, with
// the tile shapes chosen by an autotuner that times the compiled shapes on the real data.  Two launch modes:
// - single launch (the default for several classes): one launch per direction over one queue that holds the
//   items of every class, heaviest first across classes (the kernel reads each item's class from the item), with
//   one tile shape for all of them or, for the gridder, each class's own tiles per wave inside the one launch
//   (IDGX_SHAPE_TILES_PER_CLASS, the default).  The tile that suits a subgrid size is the one that covers its
//   pixels with the fewest idle lanes: S = 24 and 48 prefer fewer tiles per wave than S = 32.
// - per class (the default for one class, where the two coincide): one launch per class, each with its own tile
//   shape.
//
// Work queue.  An item is (class, subgrid, block): a block of pixels for the gridder, a chunk of visibilities
// for the degridder.  Items are built on the host from the plan (metadata, channel windows and visibility
// selections give every subgrid's visibility count V_s), costed as visibilities x pixels, and sorted heaviest
// first (longest-processing-time order), so the uneven tail (flagging, ragged timesteps) finishes together.
// Persistent workgroups (as many as fit on the chip) take items through atomic counters, which a one-workgroup
// kernel clears before the launches (no memset nodes, so the sequence can be captured).  Each queue is laid out as
// one queue per XCD: a subgrid's items all go to one of them (subgrids balanced over the queues by cost), the
// workgroups of an XCD take their own queue first and the others' when it is empty, so a subgrid's prepared pixels
// (degridder) or visibilities (gridder), which every item of the subgrid stages again, are fetched into one L2.
//
// Degridder pixels: before the persistent degridder, one workgroup per subgrid applies taper and A-terms and
// splits the pixels into the packed words the degridder stages (32 B per pixel slot), so the work items copy
// them instead of recomputing them for every chunk of visibilities.
//
// Plain FP32 batches run the persistent forms of the plain kernels (pixels per thread for the gridder, visibilities
// per thread for the degridder as their shapes), with the same queues, launch modes, capture and autotuner, and
// need no prepared pixels: their degridder items apply taper and A-terms as they stage the pixels.
//
// Device memory (queues, counters, and for the mixed forms the prepared pixels: 32 B per pixel per subgrid) is
// allocated once in idgx_batch_create; the execution entry points allocate nothing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "idgx.h"
#include "kernels.hpp"

struct idgx_batch
{
    int nr_classes = 0;
    idgx_plan plans[idgx::kMaxClasses];
    std::vector<int64_t> vis_count[idgx::kMaxClasses]; // V_s per subgrid
    int precision = IDGX_MIXED_BF16X3;
    // Per direction (0 gridder, 1 degridder) and class: tile shape, resident workgroups, queue slice.
    int shape[2][idgx::kMaxClasses] = {};
    int workgroups[2][idgx::kMaxClasses] = {};
    int nr_items[2][idgx::kMaxClasses] = {};
    size_t offset[2][idgx::kMaxClasses] = {}, capacity[2][idgx::kMaxClasses] = {};
    uint2 *d_items[2] = {};
    // Single-launch mode per direction: shape (-1 = per-class launches), item count and resident workgroups of
    // the merged queue, which lives after the class slices.
    int single_shape[2] = {-1, -1};
    int single_tiles[idgx::kMaxClasses] = {}; // gridder tiles per class of a kShapeTilesPerClass merged queue
    int single_items[2] = {};
    int last_launches[2] = {}; // persistent launches the latest run of each direction issued
    int single_workgroups[2] = {};
    size_t single_offset[2] = {};
    // Per-XCD queues: item k of queue j of a class slice (or of the merged queue) is at
    // [qbegin[j] + k]; counters [direction][class][queue] (the merged queue uses class 0's).
    int qbegin[2][idgx::kMaxClasses][idgx::kQueues + 1] = {};
    int single_qbegin[2][idgx::kQueues + 1] = {};
    int *d_counters = nullptr;
    // Degridder pixels prepared once per subgrid and pass (per class): split words, FP16 scales, slot n table.
    uint4 *d_prepared[idgx::kMaxClasses] = {};
    int *d_prepared_scale[idgx::kMaxClasses] = {};
    float4 *d_slot_coord[idgx::kMaxClasses] = {};
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

// Shape tables: the plain FP32 kernels and the mixed-precision kernels have shapes of their own.
bool plain(const idgx_batch *b)
{
    return b->precision == IDGX_PLAIN_FP32;
}

bool accurate_sincos(const idgx_batch *b)
{
    return (b->plans[0].flags & IDGX_FLAG_ACCURATE_SINCOS) != 0;
}

int shape_count(const idgx_batch *b)
{
    return plain(b) ? idgx::plain_shape_count() : idgx::mixed_shape_count();
}

int default_shape(const idgx_batch *b, bool gridder, int S)
{
    return plain(b) ? idgx::plain_default_shape(gridder, S) : idgx::mixed_default_shape(gridder, S);
}

int grid_items_per_subgrid(const idgx_batch *b, int shape, int S)
{
    return plain(b) ? idgx::plain_grid_items_per_subgrid(shape, S) : idgx::grid_items_per_subgrid(shape, S);
}

long long grid_item_pixels(const idgx_batch *b, int shape, int S, int item)
{
    return plain(b) ? idgx::plain_grid_item_pixels(shape, S, item) : idgx::grid_item_pixels(shape, S, item);
}

long long degrid_visibilities_per_item(const idgx_batch *b, int shape)
{
    return plain(b) ? idgx::plain_degrid_visibilities_per_item(shape) : idgx::degrid_visibilities_per_item(shape);
}

// Per-item tiles of a gridder single launch with each class's own shape: tiles per wave (mixed) or pixels per
// thread (plain) of the class's gridder shape.
int class_tiles(const idgx_batch *b, int shape)
{
    return plain(b) ? idgx::plain_shape(shape).pixels : idgx::mixed_shape(shape).tiles;
}

int workgroups_for(const idgx_batch *b, bool gridder, int shape)
{
    return plain(b) ? idgx::plain_queue_workgroups(accurate_sincos(b), gridder, shape)
                    : idgx::queue_workgroups(b->precision, gridder, shape);
}

hipError_t launch_queue(const idgx_batch *b, const QueueArgs &q, bool gridder, int shape, int workgroups,
                        hipStream_t st)
{
    if (plain(b))
        return gridder ? idgx::launch_grid_plain_queue(q, accurate_sincos(b), shape, workgroups, st)
                       : idgx::launch_degrid_plain_queue(q, accurate_sincos(b), shape, workgroups, st);
    return gridder ? idgx::launch_grid_mixed_queue(q, b->precision, shape, workgroups, st)
                   : idgx::launch_degrid_mixed_queue(q, b->precision, shape, workgroups, st);
}

struct Item
{
    uint2 v;
    double cost;
};

// Work items of class k with tile shape `shape`, heaviest first.
std::vector<Item> build_items(const idgx_batch *b, bool gridder, int k, int shape)
{
    std::vector<Item> items;
    const long long per = degrid_visibilities_per_item(b, shape);
    const int S = b->plans[k].subgrid_size;
    const long long P = static_cast<long long>(S) * S;
    const int gitems = grid_items_per_subgrid(b, shape, S);
    std::vector<double> gpix(gitems);
    for (int i = 0; i < gitems; ++i)
        gpix[i] = static_cast<double>(grid_item_pixels(b, shape, S, i));
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
    return items;
}

// Lays items (generated subgrid by subgrid) out as kQueues queues: each subgrid's items go to one queue, subgrids
// taken heaviest first to the least loaded queue (greedy longest-processing-time), and every queue heaviest first
// (stable: a subgrid's equal items stay together).  flat receives the queues back to back, begin their offsets.
void layout_queues(const std::vector<Item> &items, std::vector<uint2> &flat, int begin[idgx::kQueues + 1])
{
    struct Group
    {
        size_t first, count;
        double cost;
    };
    std::vector<Group> groups;
    for (size_t i = 0; i < items.size();)
    {
        size_t j = i;
        double c = 0.0;
        while (j < items.size() && items[j].v.x == items[i].v.x)
            c += items[j++].cost;
        groups.push_back({i, j - i, c});
        i = j;
    }
    std::vector<size_t> order(groups.size());
    for (size_t g = 0; g < order.size(); ++g)
        order[g] = g;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t x, size_t y) { return groups[x].cost > groups[y].cost; });
    double load[idgx::kQueues] = {};
    std::vector<Item> queue[idgx::kQueues];
    for (const size_t g : order)
    {
        int best = 0;
        for (int q = 1; q < idgx::kQueues; ++q)
            if (load[q] < load[best])
                best = q;
        load[best] += groups[g].cost;
        queue[best].insert(queue[best].end(), items.begin() + static_cast<long>(groups[g].first),
                           items.begin() + static_cast<long>(groups[g].first + groups[g].count));
    }
    flat.clear();
    flat.reserve(items.size());
    for (int q = 0; q < idgx::kQueues; ++q)
    {
        std::stable_sort(queue[q].begin(), queue[q].end(),
                         [](const Item &x, const Item &y) { return x.cost > y.cost; });
        begin[q] = static_cast<int>(flat.size());
        for (const Item &it : queue[q])
            flat.push_back(it.v);
    }
    begin[idgx::kQueues] = static_cast<int>(flat.size());
}

// Queue slice of class k: room for its items under any shape (the counts build_items gives).
size_t max_items(const idgx_batch *b, bool gridder, int k)
{
    size_t m = 0;
    for (int sh = 0; sh < shape_count(b); ++sh)
    {
        size_t n = 0;
        if (gridder)
            n = static_cast<size_t>(b->plans[k].nr_subgrids) *
                static_cast<size_t>(grid_items_per_subgrid(b, sh, b->plans[k].subgrid_size));
        else
        {
            const long long per = degrid_visibilities_per_item(b, sh);
            for (const int64_t V : b->vis_count[k])
                n += static_cast<size_t>((V + per - 1) / per);
        }
        m = std::max(m, n);
    }
    return m;
}

int upload_queue(idgx_batch *b, bool gridder, int k, int shape)
{
    const int dir = gridder ? 0 : 1;
    const std::vector<Item> items = build_items(b, gridder, k, shape);
    if (items.size() > b->capacity[dir][k] || items.size() > 0x7fffffffULL)
        return IDGX_ERR_INVALID_ARGUMENT;
    std::vector<uint2> flat;
    layout_queues(items, flat, b->qbegin[dir][k]);
    if (!flat.empty())
    {
        const hipError_t e = hipMemcpy(b->d_items[dir] + b->offset[dir][k], flat.data(), flat.size() * sizeof(uint2),
                                       hipMemcpyHostToDevice);
        if (e != hipSuccess)
            return to_status(e);
    }
    b->nr_items[dir][k] = static_cast<int>(flat.size());
    b->shape[dir][k] = shape;
    b->workgroups[dir][k] = workgroups_for(b, gridder, shape);
    return b->workgroups[dir][k] > 0 ? IDGX_OK : IDGX_ERR_DEVICE;
}

// Whether direction `dir` runs as one launch over the merged queue (a shape index or kShapeTilesPerClass).
bool single_mode(const idgx_batch *b, int dir)
{
    return b->single_shape[dir] != -1;
}

// Merged queue of every class under one shape (kShapeTilesPerClass: each class's items cut by its own gridder
// shape's tiles per wave), laid out as per-XCD queues over all classes' subgrids, heaviest first.
int upload_single_queue(idgx_batch *b, bool gridder, int shape)
{
    const int dir = gridder ? 0 : 1;
    const bool per_class = shape == idgx::kShapeTilesPerClass;
    std::vector<Item> items;
    for (int k = 0; k < b->nr_classes; ++k)
    {
        const std::vector<Item> ki = build_items(b, gridder, k, per_class ? b->shape[0][k] : shape);
        items.insert(items.end(), ki.begin(), ki.end());
        if (per_class)
            b->single_tiles[k] = class_tiles(b, b->shape[0][k]);
    }
    size_t room = 0;
    for (int k = 0; k < b->nr_classes; ++k)
        room += b->capacity[dir][k];
    if (items.size() > room || items.size() > 0x7fffffffULL)
        return IDGX_ERR_INVALID_ARGUMENT;
    std::vector<uint2> flat;
    layout_queues(items, flat, b->single_qbegin[dir]);
    if (!flat.empty())
    {
        const hipError_t e = hipMemcpy(b->d_items[dir] + b->single_offset[dir], flat.data(),
                                       flat.size() * sizeof(uint2), hipMemcpyHostToDevice);
        if (e != hipSuccess)
            return to_status(e);
    }
    b->single_items[dir] = static_cast<int>(flat.size());
    b->single_shape[dir] = shape;
    b->single_workgroups[dir] = workgroups_for(b, gridder, shape);
    return b->single_workgroups[dir] > 0 ? IDGX_OK : IDGX_ERR_DEVICE;
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
        if (!gridder && !plain(b))
        {
            a.prepared = b->d_prepared[k];
            a.prepared_scale = b->d_prepared_scale[k];
            a.slot_coord = b->d_slot_coord[k];
        }
    }
    return IDGX_OK;
}

// One persistent launch per class (each with its own tile shape), or only class `only` when it is >= 0; in
// single-launch mode (and `only` < 0) one launch over the merged queue.
int run(idgx_batch *b, const idgx_batch_data *d, bool gridder, hipStream_t st, int only = -1)
{
    const int dir = gridder ? 0 : 1;
    QueueArgs q;
    int rc = make_queue_args(b, d, gridder, q);
    if (rc != IDGX_OK)
        return rc;
    constexpr int kDirCounters = idgx::kMaxClasses * idgx::kQueues;
    hipError_t e = idgx::launch_reset_counters(b->d_counters + dir * kDirCounters, kDirCounters, st);
    b->last_launches[dir] = 0;
    if (only < 0 && single_mode(b, dir))
    {
        for (int k = 0; k < b->nr_classes && e == hipSuccess && !gridder && !plain(b); ++k)
            if (b->nr_items[dir][k] > 0)
                e = idgx::launch_degrid_prepare(q.cls[k], b->precision, st);
        if (e != hipSuccess || b->single_items[dir] == 0)
            return to_status(e);
        q.items = b->d_items[dir] + b->single_offset[dir];
        std::memcpy(q.queue_begin, b->single_qbegin[dir], sizeof(q.queue_begin));
        q.nr_queues = idgx::kQueues;
        q.counter = b->d_counters + dir * kDirCounters;
        for (int k = 0; k < idgx::kMaxClasses; ++k)
            q.tiles[k] = b->single_tiles[k];
        e = launch_queue(b, q, gridder, b->single_shape[dir], b->single_workgroups[dir], st);
        b->last_launches[dir] += e == hipSuccess;
        return to_status(e);
    }
    for (int k = 0; k < b->nr_classes && e == hipSuccess; ++k)
    {
        if ((only >= 0 && k != only) || b->nr_items[dir][k] == 0)
            continue;
        if (!gridder && !plain(b))
            e = idgx::launch_degrid_prepare(q.cls[k], b->precision, st);
        if (e != hipSuccess)
            break;
        q.items = b->d_items[dir] + b->offset[dir][k];
        std::memcpy(q.queue_begin, b->qbegin[dir][k], sizeof(q.queue_begin));
        q.nr_queues = idgx::kQueues;
        q.counter = b->d_counters + dir * kDirCounters + k * idgx::kQueues;
        e = launch_queue(b, q, gridder, b->shape[dir][k], b->workgroups[dir][k], st);
        b->last_launches[dir] += e == hipSuccess;
    }
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
            c.plan.image_size != p0.image_size || c.plan.flags != p0.flags)
            return IDGX_ERR_INVALID_ARGUMENT; // classes differ only in subgrid size / subgrids / A-term slots
        if (c.plan.nr_subgrids >= (1 << 28))
            return IDGX_ERR_UNSUPPORTED;
    }
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
    hipError_t e = hipSuccess;
    for (int dir = 0; dir < 2 && e == hipSuccess; ++dir)
    {
        size_t total = 0;
        for (int k = 0; k < nr_classes; ++k)
        {
            b->offset[dir][k] = total;
            b->capacity[dir][k] = max_items(b, dir == 0, k);
            total += b->capacity[dir][k];
        }
        // The merged queue of single-launch mode holds at most the sum of the class slices.
        b->single_offset[dir] = total;
        e = hipMalloc(&b->d_items[dir], std::max<size_t>(1, 2 * total) * sizeof(uint2));
    }
    if (e == hipSuccess)
        e = hipMalloc(&b->d_counters, 2 * idgx::kMaxClasses * idgx::kQueues * sizeof(int));
    for (int k = 0; k < nr_classes && e == hipSuccess && !plain(b); ++k)
    {
        const int S = b->plans[k].subgrid_size;
        const long long n = std::max(1, b->plans[k].nr_subgrids);
        e = hipMalloc(&b->d_prepared[k], idgx::degrid_prepared_bytes(S, n));
        if (e == hipSuccess)
            e = hipMalloc(&b->d_prepared_scale[k], n * sizeof(int));
        if (e == hipSuccess)
            e = hipMalloc(&b->d_slot_coord[k], static_cast<size_t>(S) * idgx::degrid_row_stride(S) * sizeof(float4));
    }
    int rc = to_status(e);
    for (int k = 0; k < nr_classes && rc == IDGX_OK; ++k)
    {
        const int S = b->plans[k].subgrid_size;
        rc = upload_queue(b, true, k, default_shape(b, true, S));
        if (rc == IDGX_OK)
            rc = upload_queue(b, false, k, default_shape(b, false, S));
    }
    // Several classes: the whole plan as one persistent launch per direction, gridder items with their class's
    // tiles per wave, the degridder with its default shape (the same for every subgrid size).
    if (rc == IDGX_OK && nr_classes > 1)
    {
        rc = upload_single_queue(b, true, idgx::kShapeTilesPerClass);
        if (rc == IDGX_OK)
            rc = upload_single_queue(b, false, default_shape(b, false, b->plans[0].subgrid_size));
    }
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

extern "C" int idgx_batch_plain_shape_count(void)
{
    return idgx::plain_shape_count();
}

extern "C" int idgx_batch_plain_shape_info(int shape, int *gridder_pixels_per_thread,
                                           int *degridder_visibilities_per_thread)
{
    if (shape < 0 || shape >= idgx::plain_shape_count())
        return IDGX_ERR_INVALID_ARGUMENT;
    const idgx::PlainShape s = idgx::plain_shape(shape);
    if (gridder_pixels_per_thread)
        *gridder_pixels_per_thread = s.pixels;
    if (degridder_visibilities_per_thread)
        *degridder_visibilities_per_thread = s.entries;
    return IDGX_OK;
}

extern "C" int idgx_batch_set_class_shape(idgx_batch *b, int cls, int gridder_shape, int degridder_shape)
{
    if (!b || cls < 0 || cls >= b->nr_classes || gridder_shape < 0 || gridder_shape >= shape_count(b) ||
        degridder_shape < 0 || degridder_shape >= shape_count(b))
        return IDGX_ERR_INVALID_ARGUMENT;
    drop_graph(b);
    int rc = IDGX_OK;
    if (gridder_shape != b->shape[0][cls])
    {
        rc = upload_queue(b, true, cls, gridder_shape);
        // A single launch with per-class tiles follows the classes' gridder shapes.
        if (rc == IDGX_OK && b->single_shape[0] == idgx::kShapeTilesPerClass)
            rc = upload_single_queue(b, true, idgx::kShapeTilesPerClass);
    }
    if (rc == IDGX_OK && degridder_shape != b->shape[1][cls])
        rc = upload_queue(b, false, cls, degridder_shape);
    return rc;
}

extern "C" int idgx_batch_set_shape(idgx_batch *b, int gridder_shape, int degridder_shape)
{
    if (!b)
        return IDGX_ERR_INVALID_ARGUMENT;
    int rc = IDGX_OK;
    for (int k = 0; k < b->nr_classes && rc == IDGX_OK; ++k)
        rc = idgx_batch_set_class_shape(b, k, gridder_shape, degridder_shape);
    return rc;
}

extern "C" int idgx_batch_get_class_shape(const idgx_batch *b, int cls, int *gridder_shape, int *degridder_shape)
{
    if (!b || cls < 0 || cls >= b->nr_classes)
        return IDGX_ERR_INVALID_ARGUMENT;
    if (gridder_shape)
        *gridder_shape = b->shape[0][cls];
    if (degridder_shape)
        *degridder_shape = b->shape[1][cls];
    return IDGX_OK;
}

extern "C" int idgx_batch_get_shape(const idgx_batch *b, int *gridder_shape, int *degridder_shape)
{
    return idgx_batch_get_class_shape(b, 0, gridder_shape, degridder_shape);
}

extern "C" int idgx_batch_set_single_launch(idgx_batch *b, int gridder_shape, int degridder_shape)
{
    if (!b || gridder_shape < idgx::kShapeTilesPerClass || gridder_shape >= shape_count(b) ||
        degridder_shape < -1 || degridder_shape >= shape_count(b))
        return IDGX_ERR_INVALID_ARGUMENT;
    drop_graph(b);
    const int want[2] = {gridder_shape, degridder_shape};
    int rc = IDGX_OK;
    for (int dir = 0; dir < 2 && rc == IDGX_OK; ++dir)
    {
        if (want[dir] == -1)
            b->single_shape[dir] = -1;
        else
            rc = upload_single_queue(b, dir == 0, want[dir]);
    }
    if (rc != IDGX_OK)
        b->single_shape[0] = b->single_shape[1] = -1;
    return rc;
}

extern "C" int idgx_batch_get_single_launch(const idgx_batch *b, int *gridder_shape, int *degridder_shape)
{
    if (!b)
        return IDGX_ERR_INVALID_ARGUMENT;
    if (gridder_shape)
        *gridder_shape = b->single_shape[0];
    if (degridder_shape)
        *degridder_shape = b->single_shape[1];
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
    // Average time (ms) of `repeats` runs of direction `gridder` after one warm-up (`only`: one class, or -1 for the
    // direction as it is set up); each trial ends synchronised on its event.
    const auto timed = [&](bool gridder, int only, float &ms) {
        ms = 0.F;
        int e = run(b, data, gridder, st, only); // warm-up
        if (e == IDGX_OK && hipEventRecord(e0, st) != hipSuccess)
            e = IDGX_ERR_DEVICE;
        for (int r = 0; r < repeats && e == IDGX_OK; ++r)
            e = run(b, data, gridder, st, only);
        if (e == IDGX_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                             hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            e = IDGX_ERR_DEVICE;
        ms /= repeats;
        return e;
    };
    float total[2] = {0.F, 0.F};
    // Each class is tuned on its own launch: the launches of a direction run one after another.
    for (int dir = 0; dir < 2 && rc == IDGX_OK; ++dir)
    {
        const bool gridder = dir == 0;
        const bool offer_single = single_mode(b, dir);
        for (int k = 0; k < b->nr_classes && rc == IDGX_OK; ++k)
        {
            float best = -1.F;
            int best_shape = b->shape[dir][k];
            for (int sh = 0; sh < shape_count(b) && rc == IDGX_OK; ++sh)
            {
                // One roctx range per trial, for rocprofv3 --marker-trace.
                char name[64];
                std::snprintf(name, sizeof(name), "idgx autotune %s class %d shape %d",
                              gridder ? "gridder" : "degridder", k, sh);
                roctxRangePushA(name);
                rc = upload_queue(b, gridder, k, sh);
                float ms = 0.F;
                if (rc == IDGX_OK)
                    rc = timed(gridder, k, ms);
                roctxRangePop();
                if (rc == IDGX_OK && (best < 0.F || ms < best))
                {
                    best = ms;
                    best_shape = sh;
                }
            }
            if (rc == IDGX_OK)
                rc = upload_queue(b, gridder, k, best_shape);
            total[dir] += std::max(best, 0.F);
        }
        if (rc != IDGX_OK || !offer_single)
            continue;
        // Single-launch mode: the tuned per-class launches run together (-1), then every shape of the merged queue
        // (for the gridder also each class's tuned tiles in the one launch), all timed on the whole direction; the
        // fastest of them is kept, which may leave the direction in per-class mode.
        float best = -1.F;
        int best_shape = -1;
        static_assert(idgx::kShapeTilesPerClass == -2, "the trials below run -2 (tiles per class), -1, 0, 1, ...");
        for (int sh = gridder ? idgx::kShapeTilesPerClass : -1; sh < shape_count(b) && rc == IDGX_OK; ++sh)
        {
            if (sh == -1)
                b->single_shape[dir] = -1;
            else
                rc = upload_single_queue(b, gridder, sh);
            float ms = 0.F;
            if (rc == IDGX_OK)
                rc = timed(gridder, -1, ms);
            if (rc == IDGX_OK && (best < 0.F || ms < best))
            {
                best = ms;
                best_shape = sh;
            }
        }
        if (rc == IDGX_OK && best_shape == -1)
            b->single_shape[dir] = -1;
        else if (rc == IDGX_OK)
            rc = upload_single_queue(b, gridder, best_shape);
        total[dir] = std::max(best, 0.F);
    }
    if (best_grid_ms)
        *best_grid_ms = total[0];
    if (best_degrid_ms)
        *best_degrid_ms = total[1];
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
    int items[2] = {0, 0}, wg[2] = {0, 0};
    for (int dir = 0; dir < 2; ++dir)
    {
        if (single_mode(b, dir))
        {
            items[dir] = b->single_items[dir];
            wg[dir] = b->single_workgroups[dir];
            continue;
        }
        for (int k = 0; k < b->nr_classes; ++k)
        {
            items[dir] += b->nr_items[dir][k];
            wg[dir] = std::max(wg[dir], b->workgroups[dir][k]);
        }
    }
    if (gridder_items)
        *gridder_items = items[0];
    if (degridder_items)
        *degridder_items = items[1];
    if (gridder_workgroups)
        *gridder_workgroups = wg[0];
    if (degridder_workgroups)
        *degridder_workgroups = wg[1];
    return IDGX_OK;
}

extern "C" int idgx_batch_last_launches(const idgx_batch *b, int *gridder_launches, int *degridder_launches)
{
    if (!b)
        return IDGX_ERR_INVALID_ARGUMENT;
    if (gridder_launches)
        *gridder_launches = b->last_launches[0];
    if (degridder_launches)
        *degridder_launches = b->last_launches[1];
    return IDGX_OK;
}

extern "C" void idgx_batch_destroy(idgx_batch *b)
{
    if (!b)
        return;
    drop_graph(b);
    for (int dir = 0; dir < 2; ++dir)
        if (b->d_items[dir])
            (void)hipFree(b->d_items[dir]);
    if (b->d_counters)
        (void)hipFree(b->d_counters);
    for (int k = 0; k < idgx::kMaxClasses; ++k)
    {
        if (b->d_prepared[k])
            (void)hipFree(b->d_prepared[k]);
        if (b->d_prepared_scale[k])
            (void)hipFree(b->d_prepared_scale[k]);
        if (b->d_slot_coord[k])
            (void)hipFree(b->d_slot_coord[k]);
    }
    delete b;
}
