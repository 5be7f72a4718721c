// gwn_core.hpp - shared host/device definitions of the MI355X (gfx950) geometric-query library.
//
// The public surface is the C ABI in include/gwn_hip.h; everything here is internal.
// Host code throws gwn::error; the C ABI layer translates it into a status code and a
// thread-local message, mirroring the reference's gwn_status contract
// (include/gwn/gwn_utils.cuh:213-283, include/gwn/detail/gwn_bvh_status_helpers.cuh:43-74).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <memory>
#include <vector>
#include <stdexcept>
#include <string>

#include "../../include/gwn_hip.h"
#include "../../include/gwn/detail/gwn_device_layout.hpp"

namespace gwn {

// ---------------------------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------------------------

struct error : std::runtime_error {
    int code;
    error(int c, std::string const &msg) : std::runtime_error(msg), code(c) {}
};

[[noreturn]] inline void throw_invalid(std::string const &msg) {
    throw error(GWN_ERROR_INVALID_ARGUMENT, msg);
}
[[noreturn]] inline void throw_internal(std::string const &msg) {
    throw error(GWN_ERROR_INTERNAL, msg);
}

inline void hip_check(hipError_t const e, char const *what) {
    if (e != hipSuccess) {
        throw error(
            GWN_ERROR_HIP_RUNTIME, std::string(what) + " failed [hip_error=" +
                                       std::to_string(static_cast<int>(e)) + " (" +
                                       hipGetErrorName(e) + "): \"" + hipGetErrorString(e) + "\"]"
        );
    }
}

#define GWN_HIP(call) ::gwn::hip_check((call), #call)
#define GWN_LAUNCH_CHECK() ::gwn::hip_check(hipGetLastError(), "kernel launch")

// ---------------------------------------------------------------------------------------------
// Stream-ordered device memory (hipMallocAsync pools; the reference uses cudaMallocAsync,
// include/gwn/gwn_utils.cuh:364-392).
// ---------------------------------------------------------------------------------------------

void *device_alloc(std::size_t bytes, hipStream_t stream);
void device_free(void *ptr, hipStream_t stream) noexcept;

template <class T> struct device_buffer {
    T *ptr = nullptr;
    std::size_t count = 0;
    hipStream_t stream = nullptr;

    device_buffer() = default;
    device_buffer(std::size_t n, hipStream_t s) { allocate(n, s); }
    device_buffer(device_buffer const &) = delete;
    device_buffer &operator=(device_buffer const &) = delete;
    device_buffer(device_buffer &&o) noexcept { swap(o); }
    device_buffer &operator=(device_buffer &&o) noexcept {
        device_buffer tmp(static_cast<device_buffer &&>(o));
        swap(tmp);
        return *this;
    }
    ~device_buffer() { release(); }

    void swap(device_buffer &o) noexcept {
        std::swap(ptr, o.ptr);
        std::swap(count, o.count);
        std::swap(stream, o.stream);
    }
    void allocate(std::size_t n, hipStream_t s) {
        release();
        stream = s;
        count = n;
        ptr = n ? static_cast<T *>(device_alloc(n * sizeof(T), s)) : nullptr;
    }
    void release() noexcept {
        if (ptr != nullptr)
            device_free(ptr, stream);
        ptr = nullptr;
        count = 0;
    }
    T *release_ownership() noexcept {
        T *p = ptr;
        ptr = nullptr;
        count = 0;
        return p;
    }
    std::size_t bytes() const noexcept { return count * sizeof(T); }
};

// The canonical BVH layout, triangle records and moment-record geometry are shared with the
// public device header: include/gwn/detail/gwn_device_layout.hpp.

// ---------------------------------------------------------------------------------------------
// Cross-workgroup publication within one kernel.
//
// MI355X has one L2 per XCD; agent-scope release writes dirty L2 lines back (buffer_wbl2 sc1)
// and acquire invalidates (buffer_inv sc1).  The release must also wait for the plain stores
// it publishes and for the write-back itself before the flag or counter update is performed.
// hipcc (ROCm 7.2) drops that wait when the release sits at a loop header and the stores
// were issued on the back edge, so publishers call release_agent() (fence + explicit full
// wait) and then update the flag with a relaxed atomic; consumers use acquire_agent() after
// observing it.
// ---------------------------------------------------------------------------------------------

__device__ inline void release_agent() {
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __builtin_amdgcn_s_waitcnt(0);
}

__device__ inline void acquire_agent() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

// ---------------------------------------------------------------------------------------------
// Library objects behind the opaque C handles.
// ---------------------------------------------------------------------------------------------

std::uint64_t reserve_revision() noexcept;
// Process-unique identity of one allocation of an object's device storage.  Every upload, build
// and moment refit that swaps in new buffers takes a fresh one, so holders of raw device pointers
// (batches, captured graphs) detect replaced storage even when the pool hands back the same
// addresses.
std::uint64_t next_storage_generation() noexcept;

} // namespace gwn

namespace gwn {
// Real = double storage (gwn_*_object<double, Index>).  The reference's gwn_aabb<double>,
// gwn_bvh_child<double> and gwn_bvh_node<4, double> layouts (include/gwn/gwn_bvh.cuh:25-130): a
// 48-byte box and the 64-bit child reference, 56 bytes per child.  Triangle records hold the three
// vertices, as the float records do.
struct aabb_d {
    double min_x, min_y, min_z, max_x, max_y, max_z;
};
struct bvh_child_d {
    aabb_d bounds;
    std::uint64_t reference;
};
static_assert(sizeof(bvh_child_d) == 56, "double child must be 56 bytes");
struct bvh_node4_d {
    bvh_child_d children[4];
};
struct tri_record_d {
    double v0x, v0y, v0z, v1x, v1y, v1z, v2x, v2y, v2z;
};
static_assert(sizeof(tri_record_d) == 72, "double triangle record must be 72 bytes");
} // namespace gwn

struct gwn_geometry_handle {
    gwn::device_buffer<float> x, y, z;
    gwn::device_buffer<std::uint32_t> i0, i1, i2;
    // Real = double geometry: the positions as given (xd, yd, zd); x, y, z then hold them rounded to
    // float, which is what the build's Morton codes and topology are computed from.
    gwn::device_buffer<double> xd, yd, zd;
    int real_bytes = 4;
    int index_bytes = 4; // the caller's index type (indices are stored as uint32 either way)
    std::size_t vertex_count = 0;
    std::size_t triangle_count = 0;
    hipStream_t stream = nullptr;
    std::uint64_t generation = 0;  // bumps on every upload/update
    // New on every upload and clear (which replace the buffers; an in-place update keeps it).
    std::uint64_t storage_generation = 0;
};

struct gwn_bvh_handle {
    gwn::device_buffer<gwn::bvh_node4> nodes;
    gwn::device_buffer<std::uint32_t> primitive_indices;
    gwn::device_buffer<gwn::tri_record> triangles;
    gwn::device_buffer<std::uint32_t> parent;          // (parent << 2 | slot), root = ~0u
    gwn::device_buffer<std::uint8_t> internal_children; // internal child count per node
    gwn::bvh_child root{};
    std::size_t node_count = 0;
    std::size_t triangle_count = 0;
    std::uint32_t internal_stack_bound = 0;
    std::uint32_t packed_stack_bound = 0;
    std::uint32_t max_depth = 0;
    // Children per node: 2, 3 or 4.  Nodes are stored four-wide either way; a narrower tree's
    // slots past its width stay empty, so every traversal and refit runs on it unchanged.
    int width = 4;
    // Real = double: the same tree with double boxes and records (nodes_d, triangles_d, root_d);
    // nodes, triangles and root keep the float topology build's values.
    int real_bytes = 4;
    int index_bytes = 4;
    gwn::device_buffer<gwn::bvh_node4_d> nodes_d;
    gwn::device_buffer<gwn::tri_record_d> triangles_d;
    gwn::bvh_child_d root_d{};
    std::uint64_t revision = 0;
    std::uint64_t storage_generation = 0;
    // Breadth-first levels of the wide tree: level d holds nodes [level_start[d], level_start[d + 1]);
    // the last entry is node_count.  Empty for a leaf root.
    std::vector<std::uint32_t> level_start;
    hipStream_t stream = nullptr;

    bool valid() const noexcept {
        if (revision == 0 || !gwn::ref_valid(root.reference) || triangle_count == 0)
            return false;
        if (gwn::ref_internal(root.reference))
            return gwn::ref_offset(root.reference) < node_count;
        return gwn::ref_offset(root.reference) + gwn::ref_count(root.reference) <= triangle_count;
    }
};

struct gwn_boundary_handle {
    gwn::device_buffer<std::uint32_t> start, end; // directed boundary edge endpoints
    gwn::device_buffer<std::uint64_t> multiplicity;
    std::size_t edge_count = 0;
    std::size_t mesh_vertex_count = 0;
    std::size_t mesh_triangle_count = 0;
    bool built = false;
    hipStream_t stream = nullptr;
};

struct gwn_batch_handle;

struct gwn_moment_handle {
    int order = 0;
    int width = 4;  // the width of the BVH the records were formed from
    // Real = double moments: compact child references (4 per node) and the field-SoA coefficients
    // (node, field, slot) in double.
    int real_bytes = 4;
    gwn::device_buffer<std::uint32_t> crefs_d;
    gwn::device_buffer<double> fields_d;
    gwn::device_buffer<float4> records; // node_count * moment_record_vec4(order)
    // The same count of evaluation records (taylor_eval_fields): what the value-only Taylor
    // queries read; the traversal records serve the gradient, adjoint and downloads.
    gwn::device_buffer<float4> eval_records;
    gwn::device_buffer<float> slot_area; // node_count * 4: each child slot's total triangle area
    std::size_t node_count = 0;
    std::uint64_t bvh_revision = 0;
    std::uint64_t storage_generation = 0;
    hipStream_t stream = nullptr;
};

namespace gwn {

// Level-synchronous passes over a wide tree's breadth-first levels, deepest first.  A run of
// consecutive small levels (at most `small` nodes each -- one per thread of the fused launch --
// and at most k_max_run_levels of them) goes to one single-workgroup launch that steps through
// the levels with a workgroup barrier between them (the workgroup's writes are visible to it
// after the barrier); every other level is one launch over its nodes.  Deep trees end in long
// chains of tiny levels, whose launches would otherwise cost more than their work.
constexpr int k_max_run_levels = 32;
struct level_run {
    std::uint32_t first[k_max_run_levels];
    std::uint32_t count[k_max_run_levels];
    int levels = 0;   // > 0: a fused run of small levels, in the order to process them
    std::uint32_t big_first = 0, big_count = 0;  // levels == 0: one big level
};
// The launches for levels [0, level_start.size() - 1), deepest first.
inline std::vector<level_run> plan_level_runs(std::vector<std::uint32_t> const &level_start, std::uint32_t small) {
    std::vector<level_run> runs;
    for (std::size_t d = level_start.size() - 1; d-- > 0;) {
        std::uint32_t const first = level_start[d], count = level_start[d + 1] - level_start[d];
        if (count == 0)
            continue;
        if (count <= small) {
            if (runs.empty() || runs.back().levels == 0 || runs.back().levels == k_max_run_levels)
                runs.emplace_back();
            level_run &r = runs.back();
            r.first[r.levels] = first;
            r.count[r.levels] = count;
            ++r.levels;
        } else {
            level_run r;
            r.big_first = first;
            r.big_count = count;
            runs.push_back(r);
        }
    }
    return runs;
}

// Radix sort of (key, value) pairs on bits [begin_bit, end_bit).  Keys are uint32 or uint64;
// values are uint32.  The result ends in (keys_out, vals_out); the inputs are clobbered.
void radix_sort_pairs_u64(
    std::uint64_t *keys, std::uint64_t *keys_alt, std::uint32_t *vals, std::uint32_t *vals_alt,
    std::size_t n, int begin_bit, int end_bit, std::uint64_t **keys_out,
    std::uint32_t **vals_out, hipStream_t stream
);
void radix_sort_pairs_u32(
    std::uint32_t *keys, std::uint32_t *keys_alt, std::uint32_t *vals, std::uint32_t *vals_alt,
    std::size_t n, int begin_bit, int end_bit, std::uint32_t **keys_out,
    std::uint32_t **vals_out, hipStream_t stream
);
// Device words of scratch a sort of n pairs on [begin_bit, end_bit) uses (digit bases and the
// per-tile count table).
std::size_t radix_sort_workspace_words(std::size_t n, int begin_bit, int end_bit);
// The same sorts with their scratch in `workspace` (radix_sort_workspace_words words,
// caller-owned): they allocate nothing.
void radix_sort_pairs_u64_ws(
    std::uint64_t *keys, std::uint64_t *keys_alt, std::uint32_t *vals, std::uint32_t *vals_alt,
    std::size_t n, int begin_bit, int end_bit, std::uint32_t *workspace, std::uint64_t **keys_out,
    std::uint32_t **vals_out, hipStream_t stream
);
void radix_sort_pairs_u32_ws(
    std::uint32_t *keys, std::uint32_t *keys_alt, std::uint32_t *vals, std::uint32_t *vals_alt,
    std::size_t n, int begin_bit, int end_bit, std::uint32_t *workspace, std::uint32_t **keys_out,
    std::uint32_t **vals_out, hipStream_t stream
);

// Canonical wide BVH (width 2, 3 or 4) from the LBVH or H-PLOC (search radius 1-8) binary
// hierarchy, stored four-wide.
void build_bvh4(gwn_geometry_handle const &geometry, gwn_bvh_handle &staging, int method,
                std::uint32_t hploc_radius, hipStream_t stream, int width = 4);
// Algebraic boundary of the oriented triangle chain (directed rows in (min, max) vertex order).
void build_boundary_chain(gwn_geometry_handle const &geometry, gwn_boundary_handle &staging, hipStream_t stream);
void refit_moment(gwn_bvh_handle const &bvh, gwn_moment_handle &staging, hipStream_t stream);
// Real = double (gwn_double.hip): float copies of a double geometry's positions; a BVH's double
// records, boxes and root box from the double positions over its (float-built) tree; double
// moments; the exact and Taylor winding batches in double.
void narrow_positions(gwn_geometry_handle &geometry, hipStream_t stream);
void refit_bounds_double(gwn_geometry_handle const &geometry, gwn_bvh_handle &bvh, hipStream_t stream);
void refit_moment_double(gwn_bvh_handle const &bvh, gwn_moment_handle &staging, hipStream_t stream);
void winding_exact_double(gwn_bvh_handle const &bvh, double const *qx, double const *qy, double const *qz,
                          double *out, std::size_t n, hipStream_t stream);
void winding_taylor_double(gwn_bvh_handle const &bvh, gwn_moment_handle const &m, double const *qx, double const *qy,
                           double const *qz, double *out, std::size_t n, double accuracy_scale, hipStream_t stream);
// One mesh of a batched Taylor query launch (device-resident table).
struct batch_mesh {
    float4 const *records;       // moment evaluation records (null for a leaf root)
    tri_record const *tris;
    std::uint64_t root;          // root child reference
    float const *qx, *qy, *qz;
    float *out;
    std::uint64_t count;
    std::uint64_t first_chunk;   // exclusive prefix of chunk counts over earlier meshes
    std::uint32_t const *order;  // visiting order of the queries (Hilbert), or null
    std::uint32_t wave;          // per-mesh traversal (traversal 2): 1 wave-cooperative, 0 per lane
};
// Persistent batched Taylor winding (plain precision): `grid` blocks of `block` threads take
// chunks of `chunk` queries from *counter (which must be zero) until total_chunks are done.
void taylor_batch_launch(
    batch_mesh const *meshes, std::uint32_t mesh_count, std::uint64_t total_chunks, std::uint32_t chunk,
    float accuracy_scale, std::uint64_t *counter, int order, int cap, int block, int grid, int traversal,
    hipStream_t stream
);
// Segmented builds (gwn_batch_build.hip): the H-PLOC BVH trees and Taylor moments of many meshes,
// built side by side in one pass of a fixed sequence of launches that allocates nothing and never
// reads back to the host, so the whole pass can be captured into a graph.  Each mesh's tree,
// records and moments equal those of gwn_build_bvh + gwn_refit_bvh_moment on that mesh alone.
struct batch_trees;
struct batch_trees_deleter {
    void operator()(batch_trees *t) const noexcept;
};
using batch_trees_ptr = std::unique_ptr<batch_trees, batch_trees_deleter>;
// Largest number of wide levels one pass can collapse (the level tables hold one more entry).
constexpr int k_batch_build_max_levels = 128;
// method: GWN_BVH_BUILD_HPLOC or GWN_BVH_BUILD_LBVH.
batch_trees_ptr batch_trees_create(std::vector<gwn_geometry_handle const *> const &geometries, int method,
                                   int width, std::uint32_t hploc_radius, int order, hipStream_t stream);
void batch_trees_set_build(batch_trees &t, int method, int width, std::uint32_t hploc_radius);
// Enqueues every mesh's build and moment refit, collapsing `levels` wide levels and (LBVH) bounding
// `binary_levels` binary levels at most.  Failures OR bits into the device word *status: 1 moment
// refit, 2 build, 4 a tree deeper than the launches, 8 a tree whose traversal stack bound exceeds
// `stack_capacity`.
void batch_trees_enqueue(batch_trees &t, int levels, int binary_levels, std::uint32_t stack_capacity,
                         std::uint32_t *status, hipStream_t stream);
struct batch_trees_summary {
    std::uint32_t max_depth = 0;          // deepest wide level index over all trees
    std::uint32_t max_binary_depth = 0;   // deepest LBVH binary level (LBVH builds)
    std::uint32_t max_internal_stack = 0; // largest internal_stack_bound
    std::vector<aabb> root_boxes;
};
// Synchronizes `stream` and reads the last pass's per-tree results.
batch_trees_summary batch_trees_read(batch_trees &t, hipStream_t stream);
// Mesh i's evaluation records, triangle records and root reference (fixed for the pass's life).
void batch_trees_view(batch_trees const &t, std::size_t i, float4 const **eval_records, tri_record const **tris,
                      std::uint64_t *root);
// Copies mesh i's tree and moments into single-mesh objects (allocates them; synchronizes).
void batch_trees_export(batch_trees &t, std::size_t i, gwn_bvh_handle &bvh, gwn_moment_handle &moment,
                        hipStream_t stream);
int batch_trees_width(batch_trees const &t);
int batch_trees_method(batch_trees const &t);
std::uint32_t batch_trees_radius(batch_trees const &t);
// Bytes of device memory the pass holds.
std::size_t batch_trees_bytes(batch_trees const &t);

// Batched mode (gwn_batch.hip).
gwn_batch_handle *batch_create(
    std::vector<gwn_bvh_handle const *> const &bvhs, std::vector<gwn_moment_handle *> const &moments,
    std::vector<batch_mesh> const &queries, float accuracy_scale, hipStream_t stream
);
void batch_destroy(gwn_batch_handle *b);
void batch_refit(gwn_batch_handle &b, hipStream_t stream);
void batch_query(gwn_batch_handle &b, hipStream_t stream);
// stack_capacity: 0 keeps the current compiled traversal stack, else 32, 64 or 128 (at least the
// trees' bound).
void batch_set_tuning(gwn_batch_handle &b, int block, int chunk, int blocks_per_cu, int traversal,
                      float wave_queries_per_triangle, int stack_capacity,
                      hipStream_t stream);
float batch_tune(gwn_batch_handle &b, hipStream_t stream);
void batch_get_tuning(gwn_batch_handle const &b, gwn_batch_tuning &t);
void batch_capture(gwn_batch_handle &b, bool include_refit);
void batch_replay(gwn_batch_handle &b, hipStream_t stream);
// Synchronizes `stream`, returns the batch's refit error word and clears it.
std::uint32_t batch_status(gwn_batch_handle &b, hipStream_t stream);
// The message for a nonzero status word (the most severe bit).
std::string batch_status_message(std::uint32_t bits);
// A batch over geometries: the batch builds every mesh's tree (`method` LBVH or H-PLOC with
// `hploc_radius`, `width`) and Taylor moments of `order` itself (segmented builds) and owns them;
// queries as batch_create.
gwn_batch_handle *batch_create_built(
    std::vector<gwn_geometry_handle const *> const &geometries, std::vector<batch_mesh> const &queries, int method,
    int width, std::uint32_t hploc_radius, int order, float accuracy_scale, hipStream_t stream
);
// Enqueues every mesh's build and moment refit (a batch over geometries).
void batch_build(gwn_batch_handle &b, hipStream_t stream);
// New build parameters for a batch over geometries: rebuilds now and re-plans (synchronizes).
void batch_set_build(gwn_batch_handle &b, int method, int width, std::uint32_t hploc_radius, hipStream_t stream);
void batch_get_build(gwn_batch_handle const &b, int &method, int &width, std::uint32_t &hploc_radius, int &levels,
                     int &order);
// Times builds + refits + queries over candidate builders, widths and search radii and keeps the fastest
// (a batch over geometries; synchronizes; drops a captured graph).  Returns its milliseconds.
float batch_tune_build(gwn_batch_handle &b, hipStream_t stream);
// Copies mesh i's tree and moments into BVH / moment objects (a batch over geometries).
void batch_export(gwn_batch_handle &b, std::size_t i, gwn_bvh_handle &bvh, gwn_moment_handle &moment,
                  hipStream_t stream);
// Moment refit storage that a batch allocates once per mesh, so that its refits enqueue without
// allocating or synchronising: the per-node additive payloads of the level passes.
struct moment_refit_work {
    device_buffer<std::uint8_t> pending;
};
// Allocates the moment's records and the work buffers for `bvh` (takes a new storage generation).
void refit_moment_prepare(gwn_bvh_handle const &bvh, gwn_moment_handle &m, moment_refit_work &w, hipStream_t stream);
std::size_t moment_payload_bytes(int order);
// One mesh of a batched moment refit (device table entry).
struct moment_refit_mesh {
    bvh_node4 const *nodes;
    tri_record const *tris;
    void *pending;
    float4 *records;
    float4 *eval_records;
    float *slot_area;
    std::uint64_t node_count, tri_count;
};
// One mesh's node range at one depth of a batched level pass; `prefix` counts the nodes of the
// depth's earlier segments.
struct refit_segment {
    std::uint32_t mesh;
    std::uint32_t first_node;
    std::uint64_t prefix;
};
// Segments [first_segment, first_segment + segments) of one depth, `nodes` nodes in all.
struct refit_level {
    std::uint64_t nodes;
    std::uint32_t first_segment;
    std::uint32_t segments;
};
// Every mesh's moment refit (all of one order): one launch per depth, deepest first, over all
// meshes' nodes at that depth, then one launch writing every mesh's evaluation records
// (slot_prefix: device prefix of 4 x node_count over the mesh_count meshes, plus the total).
// Records equal those of single-mesh refits; an invalid tree sets *error (device word) nonzero.
void refit_moment_batch(moment_refit_mesh const *meshes, refit_segment const *segments,
                        std::vector<refit_level> const &levels, std::uint64_t const *slot_prefix,
                        std::uint32_t mesh_count, int order, std::uint32_t *error, hipStream_t stream);
// The same with the per-depth levels (level_count of them), their segments and the slot prefix in
// device memory, written earlier in the stream (the segmented build): every launch has a fixed
// grid of `grid` blocks that strides over what the table holds, so the refit can be recorded into
// a graph before the trees it refits exist.
void refit_moment_batch_device(moment_refit_mesh const *meshes, refit_segment const *segments,
                               refit_level const *levels, int level_count, std::uint64_t const *slot_prefix,
                               std::uint32_t mesh_count, int order, std::uint32_t *error, unsigned grid,
                               hipStream_t stream);
// In-place bounds refit: triangle records and child boxes from the geometry, root box updated.
void refit_bounds(gwn_geometry_handle const &geometry, gwn_bvh_handle &bvh, hipStream_t stream);

struct query_launch {
    int block_size = 256;
    int stack_capacity = 64;
    int sort_queries = -1; // -1 auto, 0 off, 1 on; 2 (distance batch): on, wave-cooperative traversal
    int lanes_per_point = 0; // distance batch: 0 auto, 1 or 4 lanes per point (same results)
};

void winding_exact(
    gwn_bvh_handle const &bvh, float const *qx, float const *qy, float const *qz, void *out,
    std::size_t n, int precision, query_launch const &cfg, hipStream_t stream
);
// Taylor batch: winding numbers into `out` (may be null) and, when gx is non-null, their
// gradient into gx, gy, gz.
void winding_taylor(
    gwn_bvh_handle const &bvh, gwn_moment_handle const &moment, float const *qx, float const *qy,
    float const *qz, float *out, float *gx, float *gy, float *gz, std::size_t n, float accuracy_scale,
    int precision, query_launch const &cfg, std::uint32_t const *order, std::uint32_t *counters,
    hipStream_t stream
);
// Unsigned point-to-mesh distance with a culling band (queries Hilbert-sorted when cfg says so).
// `nearest` (optional, one uint32 per query) receives the BVH record of the nearest triangle, or
// k_nearest_none (distance is the band or zero: no gradient) or k_nearest_overflow (both defined
// with the device layout).
void unsigned_distance(
    gwn_bvh_handle const &bvh, float const *qx, float const *qy, float const *qz, float *out, std::size_t n,
    float culling_band, query_launch const &cfg, std::uint32_t *nearest, hipStream_t stream
);
// Work counters (2 per query: box tests, triangle tests) of the one-lane distance traversal in
// input order; distances are not written.
void unsigned_distance_work(
    gwn_bvh_handle const &bvh, float const *qx, float const *qy, float const *qz, std::size_t n, float culling_band,
    std::uint32_t *counters, hipStream_t stream
);
// Work counters (2 per ray: slab tests, triangle tests) of the batch's ray traversal.
void ray_first_hit_work(
    gwn_bvh_handle const &bvh, float const *ox, float const *oy, float const *oz, float const *dx,
    float const *dy, float const *dz, std::size_t n, float t_min, float t_max, query_launch const &cfg,
    std::uint32_t *counters, hipStream_t stream
);
// Nearest ray-triangle hits on [t_min, t_max]; `out` holds 32-byte
// gwn_ray_first_hit_result<float, uint32_t> records.
void ray_first_hit(
    gwn_bvh_handle const &bvh, float const *ox, float const *oy, float const *oz, float const *dx,
    float const *dy, float const *dz, void *out, std::size_t n, float t_min, float t_max,
    query_launch const &cfg, hipStream_t stream
);
// Adjoint traversal (internal): per-slot field adjoints (12 floats per child slot) and per-record
// near-field vertex adjoints (9 floats per record) are accumulated atomically; dq written.
void taylor_adjoint_traverse(
    gwn_bvh_handle const &bvh, gwn_moment_handle const &m, float const *qx, float const *qy, float const *qz,
    float const *u, std::uint32_t const *order, std::size_t n, float accuracy_scale, float *dslot, float *dtri,
    float *dqx, float *dqy, float *dqz, hipStream_t stream
);
// Adjoint of L = sum_q u_q w_q for the Taylor winding (orders 0, 1): adds dL/dV to dvx, dvy, dvz
// (vertex arrays of `geometry`) and writes dL/dq to dqx, dqy, dqz when they are non-null.
void winding_taylor_adjoint(
    gwn_geometry_handle const &geometry, gwn_bvh_handle const &bvh, gwn_moment_handle const &moment,
    float const *qx, float const *qy, float const *qz, float const *u, std::size_t n, float accuracy_scale,
    float *dvx, float *dvy, float *dvz, float *dqx, float *dqy, float *dqz, hipStream_t stream
);
// Adjoint of L = sum_p u_p d_p for the unsigned distance: adds dL/dV, writes dL/dp when non-null.
// With `nearest` (the forward's records) no traversal runs; otherwise the forward's traversal is
// repeated to find them.
void unsigned_distance_adjoint(
    gwn_geometry_handle const &geometry, gwn_bvh_handle const &bvh, float const *px, float const *py,
    float const *pz, float const *u, std::size_t n, float culling_band, std::uint32_t const *nearest,
    float *dvx, float *dvy, float *dvz, float *dpx, float *dpy, float *dpz, hipStream_t stream
);
// Antipodal winding numbers: positive-axis crossings over the BVH triangles plus the boundary
// chain's spherical areas; axes retried Z, X, Y, NaN when all three are singular.
void winding_antipodal(
    gwn_geometry_handle const &geometry, gwn_bvh_handle const &bvh, gwn_boundary_handle const &boundary,
    float const *qx, float const *qy, float const *qz, float *out, std::size_t n, int precision, hipStream_t stream
);
// Antipodal winding gradient (boundary term only) into gx, gy, gz.
void winding_gradient_antipodal(
    gwn_geometry_handle const &geometry, gwn_boundary_handle const &boundary, float const *qx, float const *qy,
    float const *qz, float *gx, float *gy, float *gz, std::size_t n, int precision, hipStream_t stream
);
// Hilbert-order visiting permutation of query points within the BVH root bounds.
void query_order(
    gwn_bvh_handle const &bvh, float const *qx, float const *qy, float const *qz,
    std::uint32_t *order, std::size_t n, hipStream_t stream
);
// The same within a given box (the root bounds of a tree held elsewhere).
void query_order_box(
    aabb const &box, float const *qx, float const *qy, float const *qz, std::uint32_t *order, std::size_t n,
    hipStream_t stream
);
// Taylor winding batch (no gradient) with the Hilbert presort's scratch in the library's static
// per-device scratch array instead of stream-ordered allocations; at most
// k_noalloc_max_queries queries.  Results equal winding_taylor's.
constexpr std::size_t k_noalloc_max_queries = std::size_t(1) << 20;
void winding_taylor_noalloc(
    gwn_bvh_handle const &bvh, gwn_moment_handle const &m, float const *qx, float const *qy,
    float const *qz, float *out, std::size_t n, float accuracy_scale, int precision, query_launch const &cfg,
    hipStream_t stream
);

// Zero-fill helper (a kernel, not hipMemsetAsync, so that captured graphs replay it safely).
void fill_u32(std::uint32_t *ptr, std::uint32_t value, std::size_t n, hipStream_t stream);

inline unsigned grid_for(std::size_t n, int block) {
    std::size_t g = (n + static_cast<std::size_t>(block) - 1) / static_cast<std::size_t>(block);
    return static_cast<unsigned>(g == 0 ? 1 : g);
}

} // namespace gwn
