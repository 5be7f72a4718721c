// gwn_batch.hip - batched mode: Taylor moment refits and Taylor winding queries of many meshes of
// uneven size in persistent work-queue launches, replayable as one hipGraph, with an autotuner.
//
// No reference counterpart: the reference runs one mesh per call.  A batch holds, for every mesh,
// its BVH and moment objects and its query / output device pointers:
//   - refit: every mesh's moment refit runs in one launch per tree depth over all meshes' nodes at
//     that depth, deepest first, without host synchronisation (buffers are allocated once when the
//     batch is made, so replays allocate nothing); a failed refit sets a sticky device status word
//     that gwn_batch_status reads;
//   - query: one persistent launch whose blocks take chunks of one mesh's queries from a global
//     counter (taylor_batch_kernel), so small and large meshes share the machine instead of
//     running 4096 launches back to back; every result equals the single-mesh batch's;
//   - graph: the refits, the counter reset (a fill kernel, not a memset) and the query launch are
//     captured once into a hipGraph on the batch's own stream and replayed with hipGraphLaunch;
//   - traversal: one query per lane, or wave-cooperative (a wave walks the tree once for its 64
//     queries) over each mesh's Hilbert query order, computed once when the batch is made (any
//     order gives the same results; Hilbert order keeps a wave's queries close together);
//   - tune: traversal, block size, chunk length, blocks per CU and the compiled traversal stack
//     are chosen by timing candidate launches;
//   - builds: a batch made over geometries (batch_create_built) owns every mesh's tree and
//     moments and rebuilds them all in one segmented pass (gwn_batch_build.hip) that a graph can
//     hold too, so a replay is builds + refits + queries; its tuner also picks the BVH width and
//     the H-PLOC search radius.
#include "gwn_core.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

// Device storage a batch was made over, per mesh: every build, upload and moment refit that
// swaps in new buffers takes a new storage generation.
struct gwn_batch_storage {
    std::uint64_t bvh_generation, moment_generation;
};

struct gwn_batch_handle {
    std::vector<gwn_bvh_handle const *> bvhs;
    std::vector<gwn_batch_storage> storage;
    std::vector<gwn_moment_handle *> moments;
    // A batch over geometries: the batch builds and owns the trees and moments (bvhs, moments and
    // the object refit tables stay empty).
    std::vector<gwn_geometry_handle const *> geometries;
    std::vector<std::uint64_t> geometry_storage;
    gwn::batch_trees_ptr trees;
    int build_levels = 0;          // wide levels each segmented pass collapses
    int binary_levels = 0;         // binary levels each LBVH pass bounds
    // Least traversal stack the queries may use: the trees' largest internal_stack_bound (a batch
    // over geometries adds 8 of headroom for rebuilt trees).
    std::uint32_t stack_bound = 0;
    std::vector<gwn::batch_mesh> host_meshes;
    std::vector<std::uint64_t> mesh_triangles;        // per mesh, for the per-mesh traversal rule
    std::vector<gwn::moment_refit_work> work;
    std::vector<gwn::moment_refit_mesh> host_refits;  // meshes with an internal root
    gwn::device_buffer<gwn::moment_refit_mesh> refits;
    std::vector<gwn::refit_level> refit_levels;       // per depth, deepest last
    gwn::device_buffer<gwn::refit_segment> segments;
    gwn::device_buffer<std::uint64_t> eval_prefix;    // slot prefix over host_refits (+ total)
    gwn::device_buffer<std::uint32_t> status;         // sticky refit error word
    gwn::device_buffer<gwn::batch_mesh> meshes;
    gwn::device_buffer<std::uint64_t> counter;
    std::vector<gwn::device_buffer<std::uint32_t>> orders;  // per mesh Hilbert visiting order
    std::uint64_t total_queries = 0;
    std::uint64_t total_chunks = 0;
    float accuracy_scale = 2.0f;
    int order = 1;
    int cap = 32;
    int block = 256;
    int chunk = 256;
    int blocks_per_cu = 4;
    int traversal = 0;                 // 0 per lane, 1 wave, 2 per mesh (wave_queries_per_triangle)
    float wave_queries_per_triangle = 1.0f;
    int cu_count = 1;
    hipStream_t own_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graph_refits = false;
    // Recorded after every enqueue of batch work, on the stream it went to: a re-upload of the
    // mesh table waits for it, so it never races a query or replay still reading the table.
    hipEvent_t last_work = nullptr;
    bool have_work = false;

    ~gwn_batch_handle() {
        if (last_work)
            (void)hipEventSynchronize(last_work);
        if (exec)
            (void)hipGraphExecDestroy(exec);
        if (graph)
            (void)hipGraphDestroy(graph);
        if (own_stream)
            (void)hipStreamDestroy(own_stream);
        if (last_work)
            (void)hipEventDestroy(last_work);
    }
};

namespace gwn {

void batch_drop_graph(gwn_batch_handle &b);

namespace {

void layout_chunks(gwn_batch_handle &b) {
    std::uint64_t first = 0;
    for (auto &m : b.host_meshes) {
        m.first_chunk = first;
        first += (m.count + static_cast<std::uint64_t>(b.chunk) - 1) / static_cast<std::uint64_t>(b.chunk);
    }
    b.total_chunks = first;
}

// Waits for the batch work enqueued last (queries or replays on any stream).
void wait_last_work(gwn_batch_handle &b) {
    if (b.have_work)
        GWN_HIP(hipEventSynchronize(b.last_work));
}

void note_work(gwn_batch_handle &b, hipStream_t stream) {
    GWN_HIP(hipEventRecord(b.last_work, stream));
    b.have_work = true;
}

// Traversal 2: a mesh with a query order and at least wave_queries_per_triangle queries per
// triangle walks its tree wave-cooperatively, the rest one query per lane.
void set_wave_flags(gwn_batch_handle &b) {
    for (std::size_t i = 0; i < b.host_meshes.size(); ++i) {
        batch_mesh &m = b.host_meshes[i];
        bool const dense = static_cast<double>(m.count) >=
                           static_cast<double>(b.wave_queries_per_triangle) * static_cast<double>(b.mesh_triangles[i]);
        m.wave = b.traversal == 1 || (b.traversal == 2 && m.order != nullptr && dense) ? 1u : 0u;
    }
}

void upload_table(gwn_batch_handle &b, hipStream_t stream) {
    wait_last_work(b);
    layout_chunks(b);
    set_wave_flags(b);
    GWN_HIP(hipMemcpyAsync(b.meshes.ptr, b.host_meshes.data(), b.host_meshes.size() * sizeof(batch_mesh),
                           hipMemcpyHostToDevice, stream));
    GWN_HIP(hipStreamSynchronize(stream));
}

gwn_batch_storage storage_of(gwn_bvh_handle const &bvh, gwn_moment_handle const &m) {
    return {bvh.storage_generation, m.storage_generation};
}

// Captured graphs and the device tables hold raw device pointers: refuse to run over a mesh whose
// BVH or moment storage was replaced (rebuilt, re-uploaded, refit into new buffers) after the
// batch was made.  Work that queries without refitting first also needs every moment formed
// from its BVH's current revision, as the single-mesh batches require.
void check_storage(gwn_batch_handle const &b, bool query_without_refit) {
    for (std::size_t i = 0; i < b.geometries.size(); ++i)
        if (b.geometries[i]->storage_generation != b.geometry_storage[i])
            throw_invalid("batch: A mesh's geometry storage changed since the batch was made; create a new batch.");
    for (std::size_t i = 0; i < b.bvhs.size(); ++i) {
        gwn_batch_storage const now = storage_of(*b.bvhs[i], *b.moments[i]);
        gwn_batch_storage const &was = b.storage[i];
        if (now.bvh_generation != was.bvh_generation || now.moment_generation != was.moment_generation)
            throw_invalid("batch: A mesh's BVH or moment storage changed since the batch was made; create a new batch.");
        if (b.bvhs[i]->revision == 0)
            throw_invalid("batch: A mesh's BVH is being refit; its bounds are not queryable yet.");
        if (query_without_refit && b.moments[i]->bvh_revision != b.bvhs[i]->revision)
            throw_invalid("batch: Moment object is not aligned with the BVH; refit the batch's moments first.");
    }
}

// Every mesh's moment is formed from its BVH's current bounds once enqueued refits have run.
void stamp_revisions(gwn_batch_handle &b) {
    // (A batch over geometries holds no BVH or moment objects: bvhs is empty.)
    for (std::size_t i = 0; i < b.bvhs.size(); ++i) {
        b.moments[i]->node_count = b.bvhs[i]->node_count;
        b.moments[i]->bvh_revision = b.bvhs[i]->revision;
    }
}

void enqueue_refit(gwn_batch_handle &b, hipStream_t stream) {
    if (b.trees) {
        batch_trees_enqueue(*b.trees, b.build_levels, b.binary_levels, static_cast<std::uint32_t>(b.cap), b.status.ptr,
                            stream);
        return;
    }
    refit_moment_batch(b.refits.ptr, b.segments.ptr, b.refit_levels, b.eval_prefix.ptr,
                       static_cast<std::uint32_t>(b.host_refits.size()), b.order, b.status.ptr, stream);
}

void enqueue_query(gwn_batch_handle &b, hipStream_t stream) {
    if (b.total_chunks == 0)
        return;
    fill_u32(reinterpret_cast<std::uint32_t *>(b.counter.ptr), 0u, 2, stream);
    int const grid = std::max(1, b.blocks_per_cu * b.cu_count);
    taylor_batch_launch(b.meshes.ptr, static_cast<std::uint32_t>(b.host_meshes.size()), b.total_chunks,
                        static_cast<std::uint32_t>(b.chunk), b.accuracy_scale, b.counter.ptr, b.order, b.cap,
                        b.block, grid, b.traversal, stream);
}

// The refit table: every mesh with an internal root, and for every depth the meshes' node ranges
// at that depth (segments), so one launch per depth refits all meshes.
void upload_refit_table(gwn_batch_handle &b, hipStream_t stream) {
    b.host_refits.clear();
    b.refit_levels.clear();
    std::vector<std::vector<refit_segment>> by_depth;
    for (std::size_t i = 0; i < b.bvhs.size(); ++i) {
        gwn_bvh_handle const &bvh = *b.bvhs[i];
        if (!ref_internal(bvh.root.reference) || bvh.node_count == 0)
            continue;
        auto const &ls = bvh.level_start;
        if (ls.size() < 2 || ls.front() != 0u || ls.back() != bvh.node_count)
            throw_internal("batch: A BVH level table does not cover its nodes.");
        std::uint32_t const mesh = static_cast<std::uint32_t>(b.host_refits.size());
        b.host_refits.push_back(moment_refit_mesh{bvh.nodes.ptr, bvh.triangles.ptr, b.work[i].pending.ptr,
                                                  b.moments[i]->records.ptr, b.moments[i]->eval_records.ptr,
                                                  b.moments[i]->slot_area.ptr, bvh.node_count, bvh.triangle_count});
        if (by_depth.size() < ls.size() - 1)
            by_depth.resize(ls.size() - 1);
        for (std::size_t d = 0; d + 1 < ls.size(); ++d)
            if (ls[d + 1] > ls[d])
                by_depth[d].push_back(refit_segment{mesh, ls[d], ls[d + 1] - ls[d]});  // prefix set below
    }
    std::vector<refit_segment> flat;
    for (auto &segs : by_depth) {
        refit_level lv{0, static_cast<std::uint32_t>(flat.size()), static_cast<std::uint32_t>(segs.size())};
        for (refit_segment sg : segs) {
            std::uint64_t const count = sg.prefix;
            sg.prefix = lv.nodes;
            lv.nodes += count;
            flat.push_back(sg);
        }
        b.refit_levels.push_back(lv);
    }
    std::vector<std::uint64_t> prefix(b.host_refits.size() + 1, 0);
    for (std::size_t i = 0; i < b.host_refits.size(); ++i)
        prefix[i + 1] = prefix[i] + 4 * b.host_refits[i].node_count;
    b.refits.allocate(std::max<std::size_t>(1, b.host_refits.size()), stream);
    b.segments.allocate(std::max<std::size_t>(1, flat.size()), stream);
    b.eval_prefix.allocate(prefix.size(), stream);
    GWN_HIP(hipMemcpyAsync(b.eval_prefix.ptr, prefix.data(), prefix.size() * sizeof(std::uint64_t),
                           hipMemcpyHostToDevice, stream));
    if (!b.host_refits.empty())
        GWN_HIP(hipMemcpyAsync(b.refits.ptr, b.host_refits.data(), b.host_refits.size() * sizeof(moment_refit_mesh),
                               hipMemcpyHostToDevice, stream));
    if (!flat.empty())
        GWN_HIP(hipMemcpyAsync(b.segments.ptr, flat.data(), flat.size() * sizeof(refit_segment),
                               hipMemcpyHostToDevice, stream));
    GWN_HIP(hipStreamSynchronize(stream));
}

std::uint32_t read_status(gwn_batch_handle &b, hipStream_t stream, bool clear) {
    std::uint32_t e = 0;
    GWN_HIP(hipMemcpyAsync(&e, b.status.ptr, sizeof(e), hipMemcpyDeviceToHost, stream));
    GWN_HIP(hipStreamSynchronize(stream));
    if (clear && e != 0u) {
        fill_u32(b.status.ptr, 0u, 1, stream);
        GWN_HIP(hipStreamSynchronize(stream));
    }
    return e;
}

// Smallest compiled traversal stack that holds `bound` entries.
int stack_for(std::uint32_t bound) {
    if (bound > 128)
        throw_invalid("Traversal stack bound exceeds the largest compiled stack (128).");
    return bound <= 32 ? 32 : (bound <= 64 ? 64 : 128);
}

// Builds every tree of a batch over geometries once with the largest level counts, then fixes the
// level counts of later passes and the traversal stack from the trees, each with 8 entries of
// headroom for trees rebuilt from updated geometry (synchronizes).
void plan_build(gwn_batch_handle &b, hipStream_t stream) {
    wait_last_work(b);
    batch_trees_enqueue(*b.trees, k_batch_build_max_levels, k_batch_build_max_levels, 128u, b.status.ptr, stream);
    batch_trees_summary const sum = batch_trees_read(*b.trees, stream);
    std::uint32_t const e = read_status(b, stream, true);
    if (e != 0u)
        throw_internal(batch_status_message(e));
    b.build_levels = std::min<int>(k_batch_build_max_levels, static_cast<int>(sum.max_depth) + 1 + 8);
    b.binary_levels = std::min<int>(k_batch_build_max_levels, static_cast<int>(sum.max_binary_depth) + 1 + 8);
    b.stack_bound = std::min<std::uint32_t>(128u, sum.max_internal_stack + 8u);
    b.cap = stack_for(b.stack_bound);
    batch_drop_graph(b);
}

} // namespace

// A tree cut short by the level launches (4) also leaves leaves unplaced, which sets the failed-build
// bit (2) too: the re-plan message comes first, and a build that still fails after the re-plan
// reports itself there.
std::string batch_status_message(std::uint32_t bits) {
    if (bits & 4u)
        return "batch: A rebuilt tree is deeper than the batch's level launches; call gwn_batch_set_build to re-plan.";
    if (bits & 2u)
        return "batch: A segmented BVH build failed.";
    if (bits & 8u)
        return "batch: A rebuilt tree's traversal stack bound exceeds the batch's stack capacity.";
    return "batch: Bottom-up moment refit did not converge.";
}

void batch_destroy(gwn_batch_handle *b) { delete b; }

// Every mesh's moment refit, one launch per depth over all meshes' nodes (refit_moment_batch); the
// records equal those of single-mesh refits.
void batch_refit(gwn_batch_handle &b, hipStream_t stream) {
    if (b.trees)
        throw_invalid("batch: A batch over geometries refits with its builds: call gwn_batch_build.");
    check_storage(b, false);
    enqueue_refit(b, stream);
    note_work(b, stream);
    stamp_revisions(b);
}

void batch_query(gwn_batch_handle &b, hipStream_t stream) {
    check_storage(b, true);
    enqueue_query(b, stream);
    note_work(b, stream);
}

std::uint32_t batch_status(gwn_batch_handle &b, hipStream_t stream) { return read_status(b, stream, true); }

void batch_drop_graph(gwn_batch_handle &b) {
    if (b.exec)
        GWN_HIP(hipGraphExecDestroy(b.exec));
    if (b.graph)
        GWN_HIP(hipGraphDestroy(b.graph));
    b.exec = nullptr;
    b.graph = nullptr;
}

gwn_batch_handle *batch_create(
    std::vector<gwn_bvh_handle const *> const &bvhs, std::vector<gwn_moment_handle *> const &moments,
    std::vector<batch_mesh> const &queries, float accuracy_scale, hipStream_t stream
) {
    auto *b = new gwn_batch_handle();
    try {
        GWN_HIP(hipEventCreateWithFlags(&b->last_work, hipEventDisableTiming));
        b->bvhs = bvhs;
        b->moments = moments;
        b->host_meshes = queries;
        for (gwn_bvh_handle const *h : bvhs)
            b->mesh_triangles.push_back(h->triangle_count);
        b->accuracy_scale = accuracy_scale;
        b->order = moments.empty() ? 1 : moments[0]->order;
        b->work.resize(bvhs.size());
        std::uint32_t bound = 0;
        for (std::size_t i = 0; i < bvhs.size(); ++i) {
            if (moments[i]->order != b->order)
                throw_invalid("batch: Every mesh's moment must have the same Taylor order.");
            for (std::size_t j = 0; j < i; ++j)
                if (moments[j] == moments[i])
                    throw_invalid("batch: Each mesh needs its own moment object.");
            refit_moment_prepare(*bvhs[i], *moments[i], b->work[i], stream);
            bound = std::max(bound, bvhs[i]->internal_stack_bound);
            b->storage.push_back(storage_of(*bvhs[i], *moments[i]));
        }
        b->status.allocate(1, stream);
        fill_u32(b->status.ptr, 0u, 1, stream);
        upload_refit_table(*b, stream);
        b->stack_bound = bound;
        b->cap = stack_for(bound);
        batch_refit(*b, stream);
        b->orders.resize(bvhs.size());
        for (std::size_t i = 0; i < bvhs.size(); ++i) {
            batch_mesh &m = b->host_meshes[i];
            bool const internal = ref_internal(bvhs[i]->root.reference);
            m.records = internal ? moments[i]->eval_records.ptr : nullptr;
            m.tris = bvhs[i]->triangles.ptr;
            m.root = bvhs[i]->root.reference;
            m.order = nullptr;
            if (internal && m.count > 1) {
                b->orders[i].allocate(m.count, stream);
                query_order(*bvhs[i], m.qx, m.qy, m.qz, b->orders[i].ptr, m.count, stream);
                m.order = b->orders[i].ptr;
            }
            b->total_queries += m.count;
        }
        int dev = 0;
        GWN_HIP(hipGetDevice(&dev));
        GWN_HIP(hipDeviceGetAttribute(&b->cu_count, hipDeviceAttributeMultiprocessorCount, dev));
        b->meshes.allocate(std::max<std::size_t>(1, b->host_meshes.size()), stream);
        b->counter.allocate(1, stream);
        upload_table(*b, stream);
        if (read_status(*b, stream, false) != 0u)
            throw_internal("batch: Bottom-up moment refit did not converge.");
    } catch (...) {
        delete b;
        throw;
    }
    return b;
}

void batch_set_tuning(gwn_batch_handle &b, int block, int chunk, int blocks_per_cu, int traversal,
                      float wave_queries_per_triangle, int stack_capacity, hipStream_t stream) {
    if (block <= 0 || block > 256 || (block & 63) != 0)
        throw_invalid("Batched query block size must be 64, 128, 192 or 256.");
    if (chunk <= 0 || chunk % block != 0)
        throw_invalid("Batched query chunk length must be a positive multiple of the block size.");
    if (blocks_per_cu <= 0 || blocks_per_cu > 64)
        throw_invalid("Batched query blocks per CU must be in [1, 64].");
    // 0 keeps the current stack; otherwise a compiled one (32, 64, 128) that holds every tree's bound.
    if (stack_capacity != 0 && stack_capacity != 32 && stack_capacity != 64 && stack_capacity != 128)
        throw_invalid("Batched query stack capacity must be 32, 64 or 128.");
    if (stack_capacity != 0 && static_cast<std::uint32_t>(stack_capacity) < b.stack_bound)
        throw_invalid("Batched query stack capacity is below a tree's traversal stack bound.");
    if (traversal < 0 || traversal > 2)
        throw_invalid("Batched query traversal must be 0 (per lane), 1 (wave) or 2 (per mesh).");
    if (traversal == 2 && !(wave_queries_per_triangle > 0.0f && std::isfinite(wave_queries_per_triangle)))
        throw_invalid("Batched query wave_queries_per_triangle must be positive and finite for traversal 2.");
    // The graph and any query in flight read the table and the counter: finish them first.
    wait_last_work(b);
    bool const relayout = chunk != b.chunk || traversal != b.traversal ||
                          (traversal == 2 && wave_queries_per_triangle != b.wave_queries_per_triangle);
    b.block = block;
    b.chunk = chunk;
    b.blocks_per_cu = blocks_per_cu;
    b.traversal = traversal;
    if (traversal == 2)
        b.wave_queries_per_triangle = wave_queries_per_triangle;
    if (stack_capacity != 0)
        b.cap = stack_capacity;
    if (relayout)
        upload_table(b, stream);
    batch_drop_graph(b);
}

// Times each candidate (traversal, block size, chunk length, blocks per CU) with one launch
// after a warm-up and keeps the fastest, then each compiled traversal stack that holds the trees'
// bound for that shape, then the per-mesh traversal at a few densities; returns the best time in
// milliseconds.
float batch_tune(gwn_batch_handle &b, hipStream_t stream) {
    struct cand {
        int block, chunk, bpc;
        int traversal;
        float density;
    };
    std::vector<cand> cands;
    float const density = b.wave_queries_per_triangle;
    for (int traversal : {0, 1})
        for (int block : {64, 128, 256})
            for (int mult : {1, 4})
                for (int bpc : {2, 4, 8})
                    cands.push_back({block, block * mult, bpc, traversal, density});
    check_storage(b, true);
    hipEvent_t e0, e1;
    GWN_HIP(hipEventCreate(&e0));
    GWN_HIP(hipEventCreate(&e1));
    float best_ms = 0.0f;
    cand best = {b.block, b.chunk, b.blocks_per_cu, b.traversal, density};
    bool have = false;
    try {
        for (cand const &c : cands) {
            batch_set_tuning(b, c.block, c.chunk, c.bpc, c.traversal, c.density, 0, stream);
            batch_query(b, stream);
            GWN_HIP(hipEventRecord(e0, stream));
            batch_query(b, stream);
            GWN_HIP(hipEventRecord(e1, stream));
            GWN_HIP(hipEventSynchronize(e1));
            float ms = 0.0f;
            GWN_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (!have || ms < best_ms) {
                best_ms = ms;
                best = c;
                have = true;
            }
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    batch_set_tuning(b, best.block, best.chunk, best.bpc, best.traversal, best.density, 0, stream);
    // Then the compiled traversal stack for that shape: every one that holds the trees' bound.
    int best_cap = b.cap;
    try {
        for (int cap : {32, 64, 128}) {
            if (static_cast<std::uint32_t>(cap) < b.stack_bound || cap == best_cap)
                continue;
            batch_set_tuning(b, best.block, best.chunk, best.bpc, best.traversal, best.density, cap, stream);
            batch_query(b, stream);
            GWN_HIP(hipEventRecord(e0, stream));
            batch_query(b, stream);
            GWN_HIP(hipEventRecord(e1, stream));
            GWN_HIP(hipEventSynchronize(e1));
            float ms = 0.0f;
            GWN_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best_ms) {
                best_ms = ms;
                best_cap = cap;
            }
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    // Then the per-mesh traversal (wave where a mesh has at least d queries per triangle) for that
    // shape and stack.
    try {
        for (float d : {0.0625f, 0.25f, 1.0f, 4.0f, 16.0f}) {
            batch_set_tuning(b, best.block, best.chunk, best.bpc, 2, d, best_cap, stream);
            batch_query(b, stream);
            GWN_HIP(hipEventRecord(e0, stream));
            batch_query(b, stream);
            GWN_HIP(hipEventRecord(e1, stream));
            GWN_HIP(hipEventSynchronize(e1));
            float ms = 0.0f;
            GWN_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best_ms) {
                best_ms = ms;
                best.traversal = 2;
                best.density = d;
            }
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    GWN_HIP(hipEventDestroy(e0));
    GWN_HIP(hipEventDestroy(e1));
    batch_set_tuning(b, best.block, best.chunk, best.bpc, best.traversal, best.density, best_cap, stream);
    b.wave_queries_per_triangle = best.density;  // the candidates above left their own
    return best_ms;
}

void batch_get_tuning(gwn_batch_handle const &b, gwn_batch_tuning &t) {
    t.block_size = b.block;
    t.chunk = b.chunk;
    t.blocks_per_cu = b.blocks_per_cu;
    t.traversal = b.traversal;
    t.wave_queries_per_triangle = b.wave_queries_per_triangle;
    t.stack_capacity = b.cap;
    t.meshes = b.host_meshes.size();
    t.queries = b.total_queries;
}

// Records [refits +] counter reset + query launch.  Nothing runs yet, so no moment's revision is
// stamped here: a replay with refits stamps after it is launched.
void batch_capture(gwn_batch_handle &b, bool include_refit) {
    check_storage(b, false);
    wait_last_work(b);
    batch_drop_graph(b);
    if (b.own_stream == nullptr)
        GWN_HIP(hipStreamCreateWithFlags(&b.own_stream, hipStreamNonBlocking));
    GWN_HIP(hipStreamBeginCapture(b.own_stream, hipStreamCaptureModeThreadLocal));
    try {
        if (include_refit)
            enqueue_refit(b, b.own_stream);
        enqueue_query(b, b.own_stream);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(b.own_stream, &g);
        if (g)
            (void)hipGraphDestroy(g);
        throw;
    }
    GWN_HIP(hipStreamEndCapture(b.own_stream, &b.graph));
    GWN_HIP(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
    b.graph_refits = include_refit;
}

void batch_replay(gwn_batch_handle &b, hipStream_t stream) {
    if (b.exec == nullptr)
        throw_invalid("batch: No captured graph; call gwn_batch_capture first.");
    check_storage(b, !b.graph_refits);
    GWN_HIP(hipGraphLaunch(b.exec, stream));
    note_work(b, stream);
    // The graph's refits form every moment from its BVH's current bounds, as an eager refit does.
    if (b.graph_refits)
        stamp_revisions(b);
}

// A batch over geometries: builds every tree once (plan_build fixes the level launches and the
// traversal stack), then lays out the query table over the batch's own trees and moments.
gwn_batch_handle *batch_create_built(
    std::vector<gwn_geometry_handle const *> const &geometries, std::vector<batch_mesh> const &queries, int method,
    int width, std::uint32_t hploc_radius, int order, float accuracy_scale, hipStream_t stream
) {
    auto *b = new gwn_batch_handle();
    try {
        GWN_HIP(hipEventCreateWithFlags(&b->last_work, hipEventDisableTiming));
        for (std::size_t i = 0; i < geometries.size(); ++i)
            for (std::size_t j = 0; j < i; ++j)
                if (geometries[j] == geometries[i])
                    throw_invalid("batch: Each mesh needs its own geometry object.");
        b->geometries = geometries;
        for (gwn_geometry_handle const *g : geometries)
            b->geometry_storage.push_back(g->storage_generation);
        b->host_meshes = queries;
        for (gwn_geometry_handle const *g : geometries)
            b->mesh_triangles.push_back(g->triangle_count);
        b->accuracy_scale = accuracy_scale;
        b->order = order;
        b->trees = batch_trees_create(geometries, method, width, hploc_radius, order, stream);
        b->status.allocate(1, stream);
        fill_u32(b->status.ptr, 0u, 1, stream);
        plan_build(*b, stream);
        batch_trees_summary const sum = batch_trees_read(*b->trees, stream);
        b->orders.resize(geometries.size());
        for (std::size_t i = 0; i < geometries.size(); ++i) {
            batch_mesh &m = b->host_meshes[i];
            batch_trees_view(*b->trees, i, &m.records, &m.tris, &m.root);
            m.order = nullptr;
            if (ref_internal(m.root) && m.count > 1) {
                b->orders[i].allocate(m.count, stream);
                query_order_box(sum.root_boxes[i], m.qx, m.qy, m.qz, b->orders[i].ptr, m.count, stream);
                m.order = b->orders[i].ptr;
            }
            b->total_queries += m.count;
        }
        int dev = 0;
        GWN_HIP(hipGetDevice(&dev));
        GWN_HIP(hipDeviceGetAttribute(&b->cu_count, hipDeviceAttributeMultiprocessorCount, dev));
        b->meshes.allocate(std::max<std::size_t>(1, b->host_meshes.size()), stream);
        b->counter.allocate(1, stream);
        upload_table(*b, stream);
    } catch (...) {
        delete b;
        throw;
    }
    return b;
}

void batch_build(gwn_batch_handle &b, hipStream_t stream) {
    if (!b.trees)
        throw_invalid("batch: This batch holds BVH and moment objects: refit it with gwn_batch_refit_moments.");
    check_storage(b, false);
    enqueue_refit(b, stream);
    note_work(b, stream);
}

void batch_set_build(gwn_batch_handle &b, int method, int width, std::uint32_t hploc_radius, hipStream_t stream) {
    if (!b.trees)
        throw_invalid("batch: Build parameters belong to batches over geometries.");
    check_storage(b, false);
    wait_last_work(b);
    batch_trees_set_build(*b.trees, method, width, hploc_radius);
    plan_build(b, stream);
}

void batch_get_build(gwn_batch_handle const &b, int &method, int &width, std::uint32_t &hploc_radius, int &levels,
                     int &order) {
    if (!b.trees)
        throw_invalid("batch: Build parameters belong to batches over geometries.");
    method = batch_trees_method(*b.trees);
    width = batch_trees_width(*b.trees);
    hploc_radius = batch_trees_radius(*b.trees);
    levels = b.build_levels;
    order = b.order;
}

// For every (builder, width) candidate -- LBVH, or H-PLOC with each search radius -- rebuild and
// re-plan, then time one build + refit + query pass after a warm-up one; keeps the fastest.  The
// query launch shape stays as tuned.
float batch_tune_build(gwn_batch_handle &b, hipStream_t stream) {
    if (!b.trees)
        throw_invalid("batch: Build parameters belong to batches over geometries.");
    check_storage(b, false);
    struct cand {
        int method;
        int width;
        std::uint32_t radius;
    };
    std::vector<cand> cands;
    for (int width : {4, 3, 2}) {
        cands.push_back({GWN_BVH_BUILD_LBVH, width, 8u});
        for (std::uint32_t radius : {8u, 4u, 2u, 1u})
            cands.push_back({GWN_BVH_BUILD_HPLOC, width, radius});
    }
    hipEvent_t e0, e1;
    GWN_HIP(hipEventCreate(&e0));
    GWN_HIP(hipEventCreate(&e1));
    float best_ms = 0.0f;
    cand best = {batch_trees_method(*b.trees), batch_trees_width(*b.trees), batch_trees_radius(*b.trees)};
    bool have = false;
    try {
        for (cand const &c : cands) {
            try {
                batch_set_build(b, c.method, c.width, c.radius, stream);
            } catch (error const &e) {
                if (e.code != GWN_ERROR_INTERNAL && e.code != GWN_ERROR_INVALID_ARGUMENT)
                    throw;
                continue; // trees deeper than the level limit or the compiled stacks
            }
            batch_build(b, stream);
            batch_query(b, stream);
            GWN_HIP(hipEventRecord(e0, stream));
            batch_build(b, stream);
            batch_query(b, stream);
            GWN_HIP(hipEventRecord(e1, stream));
            GWN_HIP(hipEventSynchronize(e1));
            if (read_status(b, stream, true) != 0u)
                continue; // a candidate whose trees overflow the plan is not a candidate
            float ms = 0.0f;
            GWN_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (!have || ms < best_ms) {
                best_ms = ms;
                best = c;
                have = true;
            }
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    GWN_HIP(hipEventDestroy(e0));
    GWN_HIP(hipEventDestroy(e1));
    batch_set_build(b, best.method, best.width, best.radius, stream);
    return best_ms;
}

void batch_export(gwn_batch_handle &b, std::size_t i, gwn_bvh_handle &bvh, gwn_moment_handle &moment,
                  hipStream_t stream) {
    if (!b.trees)
        throw_invalid("batch: Only a batch over geometries holds trees of its own.");
    wait_last_work(b);
    batch_trees_export(*b.trees, i, bvh, moment, stream);
}

} // namespace gwn
