// gwn_capi.hip - extern "C" surface of libgwn_hip.so (declared in include/gwn_hip.h).
//
// Every entry point validates its arguments before enqueuing work, with the reference's
// diagnostics (include/gwn/gwn_query.cuh:140-445, include/gwn/detail/gwn_bvh_build_impl.cuh:55-95,
// include/gwn/detail/gwn_geometry_impl.cuh:80-130), and translates exceptions into status codes
// the way gwn_try_translate_status does (include/gwn/detail/gwn_bvh_status_helpers.cuh:43-74).
// Builds and refits run into a staging object that replaces the target only on success (strong
// guarantee); the replaced storage is released on the stream it was allocated on.
#include "gwn_core.hpp"

#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

namespace {

thread_local std::string g_last_error;

int fail(int code, std::string const &msg) {
    g_last_error = msg;
    return code;
}

template <class F> int guarded(char const *fn, F &&f) noexcept {
    try {
        f();
        return GWN_SUCCESS;
    } catch (gwn::error const &e) {
        return fail(e.code, e.what());
    } catch (std::bad_alloc const &) {
        return fail(GWN_ERROR_INTERNAL, std::string("Host allocation failed in ") + fn + ".");
    } catch (std::exception const &e) {
        return fail(GWN_ERROR_INTERNAL, std::string("Unhandled std::exception in ") + fn + ": " + e.what());
    } catch (...) {
        return fail(GWN_ERROR_INTERNAL, std::string("Unhandled unknown exception in ") + fn + ".");
    }
}

gwn_query_config resolve_config(gwn_query_config const *cfg) {
    gwn_query_config c{256, 64, -1, 0};
    if (cfg != nullptr)
        c = *cfg;
    return c;
}

void validate_config(gwn_query_config const &c, bool traversal) {
    if (c.block_size <= 0 || c.block_size > 1024)
        gwn::throw_invalid("Query batch block size must be in the range [1, 1024].");
    if (traversal && c.stack_capacity <= 0)
        gwn::throw_invalid("Traversal stack capacity must be positive.");
}

gwn::query_launch to_launch(gwn_query_config const &c) {
    gwn::query_launch q;
    q.block_size = c.block_size;
    q.stack_capacity = c.stack_capacity;
    q.sort_queries = c.sort_queries;
    q.lanes_per_point = c.lanes_per_point;
    return q;
}

void validate_lanes(gwn_query_config const &c) {
    if (c.lanes_per_point != 0 && c.lanes_per_point != 1 && c.lanes_per_point != 4)
        gwn::throw_invalid("Lanes per point must be 0 (auto), 1 or 4.");
}

void require_bvh(gwn_bvh_handle const *bvh, int real_bytes = 4) {
    if (bvh == nullptr || !bvh->valid())
        gwn::throw_invalid("BVH object contains no queryable data.");
    if (bvh->real_bytes != real_bytes)
        gwn::throw_invalid(real_bytes == 4 ? "BVH holds Real = double: query it with the f64 entry points."
                                           : "BVH holds Real = float: query it with the f32 entry points.");
}

void require_geometry_real(gwn_geometry_handle const *g, int real_bytes) {
    if (g != nullptr && g->real_bytes != real_bytes)
        gwn::throw_invalid(real_bytes == 4 ? "Geometry holds Real = double: use the f64 entry points."
                                           : "Geometry holds Real = float: use the f32 entry points.");
}

void require_stack(gwn_bvh_handle const &bvh, int capacity) {
    if (gwn::ref_internal(bvh.root.reference) && bvh.internal_stack_bound > static_cast<std::uint32_t>(capacity))
        gwn::throw_invalid("Traversal stack capacity is below the BVH internal stack bound.");
}

void require_moment(gwn_moment_handle const *m, gwn_bvh_handle const &bvh) {
    bool ok = m != nullptr && m->bvh_revision != 0 && m->bvh_revision == bvh.revision &&
              m->real_bytes == bvh.real_bytes;
    if (ok) {
        if (gwn::ref_leaf(bvh.root.reference))
            ok = m->node_count == 0;
        else
            ok = m->node_count == bvh.node_count &&
                 (bvh.real_bytes == 8 ? m->fields_d.ptr != nullptr : m->records.ptr != nullptr);
    }
    if (!ok)
        gwn::throw_invalid("Moment object is not aligned with the BVH.");
}

void require_query_storage(std::uint64_t count, void const *a, void const *b, void const *c, void const *o) {
    if (count != 0 && (a == nullptr || b == nullptr || c == nullptr || o == nullptr))
        gwn::throw_invalid("Query/output spans must use non-null storage when non-empty.");
}

} // namespace

extern "C" {

int gwn_abi_version(void) { return GWN_HIP_ABI_VERSION; }

const char *gwn_last_error_message(void) { return g_last_error.c_str(); }

// ------------------------------------------------------------------------------------------
// geometry
// ------------------------------------------------------------------------------------------

int gwn_geometry_create(gwn_geometry_t *out) {
    return guarded("gwn_geometry_create", [&] {
        if (out == nullptr)
            gwn::throw_invalid("Output handle pointer is null.");
        *out = new gwn_geometry_handle();
    });
}

int gwn_geometry_destroy(gwn_geometry_t geometry) {
    return guarded("gwn_geometry_destroy", [&] { delete geometry; });
}

// Upload shared by every (Real, Index): Real = double keeps the positions as given and a float
// rounding of them for the topology build; uint64 indices are validated and stored as uint32
// (vertex and triangle counts stay within the 32-bit range either way).
static void upload_common(
    gwn_geometry_t geometry, int real_bytes, const void *x, const void *y, const void *z, uint64_t vertex_count,
    int index_bytes, const void *i0, const void *i1, const void *i2, uint64_t triangle_count, hipStream_t stream
) {
    if (geometry == nullptr)
        gwn::throw_invalid("Geometry handle is null.");
    if (vertex_count > 0xffffffffull)
        gwn::throw_invalid(index_bytes == 4 ? "Vertex count exceeds index type capacity."
                                            : "Vertex count exceeds the native 32-bit index range.");
    if (triangle_count > 0xffffffffull)
        gwn::throw_invalid(index_bytes == 4 ? "Triangle count exceeds index type capacity."
                                            : "Triangle count exceeds the native 32-bit index range.");
    if ((vertex_count != 0 && (x == nullptr || y == nullptr || z == nullptr)) ||
        (triangle_count != 0 && (i0 == nullptr || i1 == nullptr || i2 == nullptr)))
        gwn::throw_invalid("Geometry spans must use non-null storage when non-empty.");
    const std::uint32_t *j0 = static_cast<const std::uint32_t *>(i0);
    const std::uint32_t *j1 = static_cast<const std::uint32_t *>(i1);
    const std::uint32_t *j2 = static_cast<const std::uint32_t *>(i2);
    std::vector<std::uint32_t> narrowed;
    if (index_bytes == 8) {
        auto const *k0 = static_cast<const std::uint64_t *>(i0);
        auto const *k1 = static_cast<const std::uint64_t *>(i1);
        auto const *k2 = static_cast<const std::uint64_t *>(i2);
        for (std::uint64_t t = 0; t < triangle_count; ++t)
            if (k0[t] >= vertex_count || k1[t] >= vertex_count || k2[t] >= vertex_count)
                gwn::throw_invalid("Triangle indices must be in [0, vertex_count).");
        narrowed.resize(3 * triangle_count);
        for (std::uint64_t t = 0; t < triangle_count; ++t) {
            narrowed[t] = static_cast<std::uint32_t>(k0[t]);
            narrowed[triangle_count + t] = static_cast<std::uint32_t>(k1[t]);
            narrowed[2 * triangle_count + t] = static_cast<std::uint32_t>(k2[t]);
        }
        j0 = narrowed.data();
        j1 = j0 + triangle_count;
        j2 = j1 + triangle_count;
    } else {
        for (std::uint64_t t = 0; t < triangle_count; ++t)
            if (j0[t] >= vertex_count || j1[t] >= vertex_count || j2[t] >= vertex_count)
                gwn::throw_invalid("Triangle indices must be in [0, vertex_count).");
    }
    gwn_geometry_handle staging;
    staging.stream = stream;
    staging.vertex_count = vertex_count;
    staging.triangle_count = triangle_count;
    staging.real_bytes = real_bytes;
    staging.index_bytes = index_bytes;
    staging.x.allocate(vertex_count, stream);
    staging.y.allocate(vertex_count, stream);
    staging.z.allocate(vertex_count, stream);
    staging.i0.allocate(triangle_count, stream);
    staging.i1.allocate(triangle_count, stream);
    staging.i2.allocate(triangle_count, stream);
    std::size_t const tb = triangle_count * sizeof(std::uint32_t);
    if (real_bytes == 8) {
        staging.xd.allocate(vertex_count, stream);
        staging.yd.allocate(vertex_count, stream);
        staging.zd.allocate(vertex_count, stream);
        std::size_t const vb = vertex_count * sizeof(double);
        if (vb) {
            GWN_HIP(hipMemcpyAsync(staging.xd.ptr, x, vb, hipMemcpyHostToDevice, stream));
            GWN_HIP(hipMemcpyAsync(staging.yd.ptr, y, vb, hipMemcpyHostToDevice, stream));
            GWN_HIP(hipMemcpyAsync(staging.zd.ptr, z, vb, hipMemcpyHostToDevice, stream));
        }
        gwn::narrow_positions(staging, stream);
    } else {
        std::size_t const vb = vertex_count * sizeof(float);
        if (vb) {
            GWN_HIP(hipMemcpyAsync(staging.x.ptr, x, vb, hipMemcpyHostToDevice, stream));
            GWN_HIP(hipMemcpyAsync(staging.y.ptr, y, vb, hipMemcpyHostToDevice, stream));
            GWN_HIP(hipMemcpyAsync(staging.z.ptr, z, vb, hipMemcpyHostToDevice, stream));
        }
    }
    if (tb) {
        GWN_HIP(hipMemcpyAsync(staging.i0.ptr, j0, tb, hipMemcpyHostToDevice, stream));
        GWN_HIP(hipMemcpyAsync(staging.i1.ptr, j1, tb, hipMemcpyHostToDevice, stream));
        GWN_HIP(hipMemcpyAsync(staging.i2.ptr, j2, tb, hipMemcpyHostToDevice, stream));
    }
    // Host sources may be pageable and released by the caller after return.
    GWN_HIP(hipStreamSynchronize(stream));
    staging.generation = geometry->generation + 1;
    std::swap(staging.x, geometry->x);
    std::swap(staging.y, geometry->y);
    std::swap(staging.z, geometry->z);
    std::swap(staging.xd, geometry->xd);
    std::swap(staging.yd, geometry->yd);
    std::swap(staging.zd, geometry->zd);
    std::swap(staging.i0, geometry->i0);
    std::swap(staging.i1, geometry->i1);
    std::swap(staging.i2, geometry->i2);
    geometry->vertex_count = staging.vertex_count;
    geometry->triangle_count = staging.triangle_count;
    geometry->real_bytes = real_bytes;
    geometry->index_bytes = index_bytes;
    geometry->stream = stream;
    geometry->generation = staging.generation;
    geometry->storage_generation = gwn::next_storage_generation();
}

int gwn_upload_geometry_f32(
    gwn_geometry_t geometry, const float *x, const float *y, const float *z, uint64_t vertex_count,
    const uint32_t *i0, const uint32_t *i1, const uint32_t *i2, uint64_t triangle_count, hipStream_t stream
) {
    return guarded("gwn_upload_geometry", [&] {
        upload_common(geometry, 4, x, y, z, vertex_count, 4, i0, i1, i2, triangle_count, stream);
    });
}

int gwn_upload_geometry_f32_u64(
    gwn_geometry_t geometry, const float *x, const float *y, const float *z, uint64_t vertex_count,
    const uint64_t *i0, const uint64_t *i1, const uint64_t *i2, uint64_t triangle_count, hipStream_t stream
) {
    return guarded("gwn_upload_geometry", [&] {
        upload_common(geometry, 4, x, y, z, vertex_count, 8, i0, i1, i2, triangle_count, stream);
    });
}

int gwn_upload_geometry_f64(
    gwn_geometry_t geometry, const double *x, const double *y, const double *z, uint64_t vertex_count,
    const uint32_t *i0, const uint32_t *i1, const uint32_t *i2, uint64_t triangle_count, hipStream_t stream
) {
    return guarded("gwn_upload_geometry", [&] {
        upload_common(geometry, 8, x, y, z, vertex_count, 4, i0, i1, i2, triangle_count, stream);
    });
}

int gwn_upload_geometry_f64_u64(
    gwn_geometry_t geometry, const double *x, const double *y, const double *z, uint64_t vertex_count,
    const uint64_t *i0, const uint64_t *i1, const uint64_t *i2, uint64_t triangle_count, hipStream_t stream
) {
    return guarded("gwn_upload_geometry", [&] {
        upload_common(geometry, 8, x, y, z, vertex_count, 8, i0, i1, i2, triangle_count, stream);
    });
}

// Object stream binding (include/gwn/gwn_utils.cuh:59-80): every buffer of an object is released
// on the object's bound stream, so rebinding moves every buffer's release stream with it.
extern "C++" {
static void rebind(gwn_geometry_handle &g, hipStream_t s) noexcept {
    g.stream = s;
    g.x.stream = g.y.stream = g.z.stream = s;
    g.i0.stream = g.i1.stream = g.i2.stream = s;
    g.xd.stream = g.yd.stream = g.zd.stream = s;
}

static void rebind(gwn_bvh_handle &b, hipStream_t s) noexcept {
    b.stream = s;
    b.nodes.stream = b.primitive_indices.stream = b.triangles.stream = s;
    b.parent.stream = b.internal_children.stream = s;
    b.nodes_d.stream = b.triangles_d.stream = s;
}

static void rebind(gwn_moment_handle &m, hipStream_t s) noexcept {
    m.stream = s;
    m.records.stream = m.eval_records.stream = m.slot_area.stream = s;
    m.crefs_d.stream = m.fields_d.stream = s;
}

static void rebind(gwn_boundary_handle &b, hipStream_t s) noexcept {
    b.stream = s;
    b.start.stream = b.end.stream = b.multiplicity.stream = s;
}
} // extern "C++"

static void update_positions(
    gwn_geometry_t geometry, int real_bytes, const void *x, const void *y, const void *z, uint64_t vertex_count,
    hipStream_t stream, hipMemcpyKind kind
) {
    if (geometry == nullptr || (geometry->vertex_count != 0 && geometry->x.ptr == nullptr))
        gwn::throw_invalid("Geometry accessor is invalid for update.");
    if (geometry->real_bytes != real_bytes)
        gwn::throw_invalid(real_bytes == 4 ? "Geometry holds Real = double: update it with the f64 entry points."
                                           : "Geometry holds Real = float: update it with the f32 entry points.");
    if (vertex_count != geometry->vertex_count)
        gwn::throw_invalid("Updated vertex count must match geometry.");
    if (vertex_count != 0 && (x == nullptr || y == nullptr || z == nullptr))
        gwn::throw_invalid("Geometry update spans must use non-null storage when non-empty.");
    // From the first copy on, the storage holds this stream's writes: release it there.
    rebind(*geometry, stream);
    std::size_t const vb = vertex_count * static_cast<std::size_t>(real_bytes);
    if (vb) {
        if (real_bytes == 8) {
            GWN_HIP(hipMemcpyAsync(geometry->xd.ptr, x, vb, kind, stream));
            GWN_HIP(hipMemcpyAsync(geometry->yd.ptr, y, vb, kind, stream));
            GWN_HIP(hipMemcpyAsync(geometry->zd.ptr, z, vb, kind, stream));
            gwn::narrow_positions(*geometry, stream);
        } else {
            GWN_HIP(hipMemcpyAsync(geometry->x.ptr, x, vb, kind, stream));
            GWN_HIP(hipMemcpyAsync(geometry->y.ptr, y, vb, kind, stream));
            GWN_HIP(hipMemcpyAsync(geometry->z.ptr, z, vb, kind, stream));
        }
    }
    geometry->generation += 1;
}

int gwn_update_geometry_f32(
    gwn_geometry_t geometry, const float *x, const float *y, const float *z, uint64_t vertex_count, hipStream_t stream
) {
    return guarded("gwn_update_geometry", [&] {
        // Enqueued like the reference's update (detail/gwn_geometry_impl.cuh:172-178): from pageable
        // host memory the runtime stages the bytes before the call returns; pinned sources must
        // stay valid until the stream reaches the copies.
        update_positions(geometry, 4, x, y, z, vertex_count, stream, hipMemcpyHostToDevice);
    });
}

int gwn_update_geometry_device_f32(
    gwn_geometry_t geometry, const float *x, const float *y, const float *z, uint64_t vertex_count, hipStream_t stream
) {
    return guarded("gwn_update_geometry", [&] {
        update_positions(geometry, 4, x, y, z, vertex_count, stream, hipMemcpyDeviceToDevice);
    });
}

int gwn_update_geometry_f64(
    gwn_geometry_t geometry, const double *x, const double *y, const double *z, uint64_t vertex_count,
    hipStream_t stream
) {
    return guarded("gwn_update_geometry", [&] {
        update_positions(geometry, 8, x, y, z, vertex_count, stream, hipMemcpyHostToDevice);
    });
}

int gwn_update_geometry_device_f64(
    gwn_geometry_t geometry, const double *x, const double *y, const double *z, uint64_t vertex_count,
    hipStream_t stream
) {
    return guarded("gwn_update_geometry", [&] {
        update_positions(geometry, 8, x, y, z, vertex_count, stream, hipMemcpyDeviceToDevice);
    });
}

int gwn_geometry_counts(gwn_geometry_t geometry, uint64_t *vertex_count, uint64_t *triangle_count) {
    return guarded("gwn_geometry_counts", [&] {
        if (geometry == nullptr)
            gwn::throw_invalid("Geometry handle is null.");
        if (vertex_count)
            *vertex_count = geometry->vertex_count;
        if (triangle_count)
            *triangle_count = geometry->triangle_count;
    });
}

// ------------------------------------------------------------------------------------------
// BVH
// ------------------------------------------------------------------------------------------

int gwn_bvh_create(gwn_bvh_t *out) {
    return guarded("gwn_bvh_create", [&] {
        if (out == nullptr)
            gwn::throw_invalid("Output handle pointer is null.");
        *out = new gwn_bvh_handle();
    });
}

int gwn_bvh_destroy(gwn_bvh_t bvh) {
    return guarded("gwn_bvh_destroy", [&] { delete bvh; });
}

static void build_common(
    gwn_geometry_t geometry, gwn_bvh_t bvh, int32_t width, int32_t method, uint32_t hploc_search_radius,
    hipStream_t stream, int real_bytes = 4
);

int gwn_build_bvh4_f32(
    gwn_geometry_t geometry, gwn_bvh_t bvh, int32_t method, uint32_t hploc_search_radius, hipStream_t stream
) {
    return guarded("gwn_build_bvh", [&] { build_common(geometry, bvh, 4, method, hploc_search_radius, stream); });
}

int gwn_build_bvh_f32(
    gwn_geometry_t geometry, gwn_bvh_t bvh, int32_t width, int32_t method, uint32_t hploc_search_radius,
    hipStream_t stream
) {
    return guarded("gwn_build_bvh", [&] { build_common(geometry, bvh, width, method, hploc_search_radius, stream); });
}

int gwn_build_bvh_f64(
    gwn_geometry_t geometry, gwn_bvh_t bvh, int32_t width, int32_t method, uint32_t hploc_search_radius,
    hipStream_t stream
) {
    return guarded("gwn_build_bvh", [&] {
        build_common(geometry, bvh, width, method, hploc_search_radius, stream, 8);
    });
}

static void build_common(
    gwn_geometry_t geometry, gwn_bvh_t bvh, int32_t width, int32_t method, uint32_t hploc_search_radius,
    hipStream_t stream, int real_bytes
) {
    {
        if (bvh == nullptr)
            gwn::throw_invalid("bvh.build: BVH handle is null.");
        if (geometry == nullptr || (geometry->triangle_count != 0 && geometry->i0.ptr == nullptr))
            gwn::throw_invalid("bvh.build: Geometry accessor is invalid for BVH build.");
        if (method != GWN_BVH_BUILD_HPLOC && method != GWN_BVH_BUILD_LBVH)
            gwn::throw_invalid("bvh.build: BVH build method is invalid.");
        if (width < 2 || width > 4)
            gwn::throw_invalid("bvh.build: BVH width must be 2, 3 or 4 (nodes are stored four-wide).");
        if (method == GWN_BVH_BUILD_HPLOC && (hploc_search_radius == 0 || hploc_search_radius > 8))
            gwn::throw_invalid("bvh.build: H-PLOC search radius must be in the range [1, 8].");
        if (geometry->triangle_count != 0 && geometry->vertex_count == 0)
            gwn::throw_invalid("bvh.build: Cannot build a BVH with triangles but no vertices.");
        require_geometry_real(geometry, real_bytes);
        // Wide-node ids and per-level counts travel in 30-bit fields (look-back status words,
        // parent links with the slot in the low two bits).
        if (geometry->triangle_count > (1ull << 30) - 1ull)
            gwn::throw_invalid("bvh.build: Triangle count exceeds the native wide-collapse counter range.");
        gwn_bvh_handle staging;
        staging.stream = stream;
        gwn::build_bvh4(*geometry, staging, method, hploc_search_radius, stream, width);
        staging.real_bytes = real_bytes;
        staging.index_bytes = geometry->index_bytes;
        // Real = double: double records and boxes from the given positions over the float-built tree.
        if (real_bytes == 8 && geometry->triangle_count != 0)
            gwn::refit_bounds_double(*geometry, staging, stream);
        staging.storage_generation = gwn::next_storage_generation();
        if (geometry->triangle_count != 0) {
            staging.revision = gwn::reserve_revision();
            if (staging.revision == 0)
                gwn::throw_internal("bvh.build: BVH geometry-derived state revision is exhausted.");
        }
        std::swap(*bvh, staging);
        // staging now owns the replaced storage; its buffers free on their own streams.
    }
}

static void refit_common(gwn_geometry_t geometry, gwn_bvh_t bvh, hipStream_t stream, int real_bytes) {
    if (geometry == nullptr || (geometry->triangle_count != 0 && geometry->i0.ptr == nullptr))
        gwn::throw_invalid("bvh.refit: Geometry accessor is invalid.");
    if (bvh == nullptr || !bvh->valid())
        gwn::throw_invalid("bvh.refit: BVH accessor is invalid.");
    require_geometry_real(geometry, real_bytes);
    if (bvh->real_bytes != real_bytes)
        gwn::throw_invalid("bvh.refit: BVH and refit entry point differ in Real.");
    if (geometry->triangle_count != bvh->triangle_count)
        gwn::throw_invalid("bvh.refit: Geometry triangle count differs from the BVH primitive order.");
    std::uint64_t const next_revision = gwn::reserve_revision();
    if (next_revision == 0)
        gwn::throw_internal("bvh.refit: BVH geometry-derived state revision is exhausted.");
    // Writes below are in place and cannot be rolled back: the BVH is unqueryable (and
    // every moment formed from it stale) until the refit has converged.  From here on its
    // storage is released on the refit stream, which holds the last writes.
    bvh->revision = 0;
    rebind(*bvh, stream);
    gwn::refit_bounds(*geometry, *bvh, stream);
    if (real_bytes == 8)
        gwn::refit_bounds_double(*geometry, *bvh, stream);
    bvh->revision = next_revision;
}

int gwn_refit_bvh_f32(gwn_geometry_t geometry, gwn_bvh_t bvh, hipStream_t stream) {
    return guarded("gwn_refit_bvh", [&] { refit_common(geometry, bvh, stream, 4); });
}

int gwn_refit_bvh_f64(gwn_geometry_t geometry, gwn_bvh_t bvh, hipStream_t stream) {
    return guarded("gwn_refit_bvh", [&] { refit_common(geometry, bvh, stream, 8); });
}

int gwn_bvh_get_info(gwn_bvh_t bvh, gwn_bvh_info *info) {
    return guarded("gwn_bvh_get_info", [&] {
        if (bvh == nullptr || info == nullptr)
            gwn::throw_invalid("BVH handle or info pointer is null.");
        std::memset(info, 0, sizeof(*info));
        info->node_count = bvh->node_count;
        info->triangle_count = bvh->revision ? bvh->triangle_count : 0;
        info->revision = bvh->revision;
        info->internal_stack_bound = bvh->internal_stack_bound;
        info->packed_stack_bound = bvh->packed_stack_bound;
        info->max_depth = bvh->max_depth;
        info->width = static_cast<uint32_t>(bvh->width);
        std::memcpy(info->root_bounds, &bvh->root.bounds, sizeof(info->root_bounds));
        info->root_reference = bvh->root.reference;
    });
}

int gwn_bvh_download_f64(gwn_bvh_t bvh, void *nodes, uint32_t *primitive_indices, double *triangles,
                         hipStream_t stream) {
    return guarded("gwn_bvh_download", [&] {
        require_bvh(bvh, 8);
        std::vector<gwn::bvh_node4_d> wide;
        if (nodes && bvh->node_count) {
            wide.resize(bvh->node_count);
            GWN_HIP(hipMemcpyAsync(wide.data(), bvh->nodes_d.ptr, bvh->node_count * sizeof(gwn::bvh_node4_d),
                                   hipMemcpyDeviceToHost, stream));
        }
        if (primitive_indices)
            GWN_HIP(hipMemcpyAsync(primitive_indices, bvh->primitive_indices.ptr,
                                   bvh->triangle_count * sizeof(std::uint32_t), hipMemcpyDeviceToHost, stream));
        if (triangles)
            GWN_HIP(hipMemcpyAsync(triangles, bvh->triangles_d.ptr, bvh->triangle_count * sizeof(gwn::tri_record_d),
                                   hipMemcpyDeviceToHost, stream));
        GWN_HIP(hipStreamSynchronize(stream));
        // gwn_bvh_node<Width, double>: each node's first Width children.
        if (!wide.empty()) {
            std::size_t const child_bytes = sizeof(gwn::bvh_child_d);
            auto *dst = static_cast<std::uint8_t *>(nodes);
            for (std::size_t n = 0; n < wide.size(); ++n)
                std::memcpy(dst + n * bvh->width * child_bytes, wide[n].children, bvh->width * child_bytes);
        }
    });
}

int gwn_bvh_download(gwn_bvh_t bvh, void *nodes, uint32_t *primitive_indices, float *triangles, hipStream_t stream) {
    return guarded("gwn_bvh_download", [&] {
        require_bvh(bvh);
        std::vector<gwn::bvh_node4> wide;
        if (nodes && bvh->node_count) {
            if (bvh->width == 4) {
                GWN_HIP(hipMemcpyAsync(nodes, bvh->nodes.ptr, bvh->node_count * sizeof(gwn::bvh_node4),
                                       hipMemcpyDeviceToHost, stream));
            } else {
                wide.resize(bvh->node_count);
                GWN_HIP(hipMemcpyAsync(wide.data(), bvh->nodes.ptr, bvh->node_count * sizeof(gwn::bvh_node4),
                                       hipMemcpyDeviceToHost, stream));
            }
        }
        if (primitive_indices)
            GWN_HIP(hipMemcpyAsync(primitive_indices, bvh->primitive_indices.ptr, bvh->triangle_count * sizeof(std::uint32_t), hipMemcpyDeviceToHost, stream));
        if (triangles)
            GWN_HIP(hipMemcpyAsync(triangles, bvh->triangles.ptr, bvh->triangle_count * sizeof(gwn::tri_record), hipMemcpyDeviceToHost, stream));
        GWN_HIP(hipStreamSynchronize(stream));
        // A narrower tree is the reference's gwn_bvh_node<Width, float>: its first Width children.
        if (!wide.empty()) {
            std::size_t const child_bytes = sizeof(gwn::bvh_child);
            auto *dst = static_cast<std::uint8_t *>(nodes);
            for (std::size_t n = 0; n < wide.size(); ++n)
                std::memcpy(dst + n * bvh->width * child_bytes, wide[n].children, bvh->width * child_bytes);
        }
    });
}

// ------------------------------------------------------------------------------------------
// object lifetime: stream binding, release, accessors
// ------------------------------------------------------------------------------------------

extern "C++" template <class H> static H *require_handle(H *h, char const *what) {
    if (h == nullptr)
        gwn::throw_invalid(std::string(what) + " handle is null.");
    return h;
}

int gwn_geometry_stream(gwn_geometry_t geometry, hipStream_t *stream) {
    return guarded("gwn_geometry_stream", [&] {
        if (stream == nullptr)
            gwn::throw_invalid("Output stream pointer is null.");
        *stream = require_handle(geometry, "Geometry")->stream;
    });
}

int gwn_geometry_set_stream(gwn_geometry_t geometry, hipStream_t stream) {
    return guarded("gwn_geometry_set_stream", [&] { rebind(*require_handle(geometry, "Geometry"), stream); });
}

int gwn_geometry_clear(gwn_geometry_t geometry) {
    return guarded("gwn_geometry_clear", [&] {
        gwn_geometry_handle &g = *require_handle(geometry, "Geometry");
        for (auto *b : {&g.x, &g.y, &g.z})
            b->release();
        for (auto *b : {&g.xd, &g.yd, &g.zd})
            b->release();
        for (auto *b : {&g.i0, &g.i1, &g.i2})
            b->release();
        g.vertex_count = g.triangle_count = 0;
        g.generation += 1;
        g.storage_generation = gwn::next_storage_generation();
    });
}

int gwn_geometry_get_accessor(gwn_geometry_t geometry, gwn_geometry_accessor *out) {
    return guarded("gwn_geometry_get_accessor", [&] {
        gwn_geometry_handle const &g = *require_handle(geometry, "Geometry");
        if (out == nullptr)
            gwn::throw_invalid("Accessor output pointer is null.");
        if (g.real_bytes != 4)
            gwn::throw_invalid("Device accessors are provided for Real = float objects.");
        *out = gwn_geometry_accessor{g.x.ptr, g.y.ptr, g.z.ptr, g.i0.ptr, g.i1.ptr, g.i2.ptr,
                                     g.vertex_count, g.triangle_count};
    });
}

int gwn_bvh_stream(gwn_bvh_t bvh, hipStream_t *stream) {
    return guarded("gwn_bvh_stream", [&] {
        if (stream == nullptr)
            gwn::throw_invalid("Output stream pointer is null.");
        *stream = require_handle(bvh, "BVH")->stream;
    });
}

int gwn_bvh_set_stream(gwn_bvh_t bvh, hipStream_t stream) {
    return guarded("gwn_bvh_set_stream", [&] { rebind(*require_handle(bvh, "BVH"), stream); });
}

int gwn_bvh_clear(gwn_bvh_t bvh) {
    return guarded("gwn_bvh_clear", [&] {
        gwn_bvh_handle &b = *require_handle(bvh, "BVH");
        hipStream_t const s = b.stream;
        gwn_bvh_handle empty;
        std::swap(b, empty);
        rebind(b, s);
        b.storage_generation = gwn::next_storage_generation();
        // `empty` releases the old storage on the old binding as it goes out of scope.
    });
}

int gwn_bvh_get_accessor(gwn_bvh_t bvh, gwn_bvh_accessor *out) {
    return guarded("gwn_bvh_get_accessor", [&] {
        gwn_bvh_handle const &b = *require_handle(bvh, "BVH");
        if (out == nullptr)
            gwn::throw_invalid("Accessor output pointer is null.");
        if (b.real_bytes != 4)
            gwn::throw_invalid("Device accessors are provided for Real = float objects.");
        std::memset(out, 0, sizeof(*out));
        out->nodes = b.nodes.ptr;
        out->node_count = b.node_count;
        out->primitive_indices = b.primitive_indices.ptr;
        out->triangles = reinterpret_cast<const float *>(b.triangles.ptr);
        out->triangle_count = b.triangle_count;
        std::memcpy(out->root_bounds, &b.root.bounds, sizeof(out->root_bounds));
        out->root_reference = b.root.reference;
        out->internal_stack_bound = b.internal_stack_bound;
        out->packed_stack_bound = b.packed_stack_bound;
        out->revision = b.revision;
    });
}

int gwn_moment_stream(gwn_moment_t moment, hipStream_t *stream) {
    return guarded("gwn_moment_stream", [&] {
        if (stream == nullptr)
            gwn::throw_invalid("Output stream pointer is null.");
        *stream = require_handle(moment, "Moment")->stream;
    });
}

int gwn_moment_set_stream(gwn_moment_t moment, hipStream_t stream) {
    return guarded("gwn_moment_set_stream", [&] { rebind(*require_handle(moment, "Moment"), stream); });
}

int gwn_moment_clear(gwn_moment_t moment) {
    return guarded("gwn_moment_clear", [&] {
        gwn_moment_handle &m = *require_handle(moment, "Moment");
        m.records.release();
        m.eval_records.release();
        m.slot_area.release();
        m.crefs_d.release();
        m.fields_d.release();
        m.node_count = 0;
        m.bvh_revision = 0;
        m.storage_generation = gwn::next_storage_generation();
    });
}

int gwn_moment_get_accessor(gwn_moment_t moment, gwn_moment_accessor *out) {
    return guarded("gwn_moment_get_accessor", [&] {
        gwn_moment_handle const &m = *require_handle(moment, "Moment");
        if (out == nullptr)
            gwn::throw_invalid("Accessor output pointer is null.");
        if (m.real_bytes != 4)
            gwn::throw_invalid("Device accessors are provided for Real = float objects.");
        *out = gwn_moment_accessor{m.records.ptr, m.node_count, m.order,
                                   static_cast<uint32_t>(gwn::moment_record_vec4(m.order) * 16), m.width,
                                   m.bvh_revision, m.eval_records.ptr};
    });
}

int gwn_boundary_stream(gwn_boundary_t boundary, hipStream_t *stream) {
    return guarded("gwn_boundary_stream", [&] {
        if (stream == nullptr)
            gwn::throw_invalid("Output stream pointer is null.");
        *stream = require_handle(boundary, "Boundary chain")->stream;
    });
}

int gwn_boundary_set_stream(gwn_boundary_t boundary, hipStream_t stream) {
    return guarded("gwn_boundary_set_stream", [&] { rebind(*require_handle(boundary, "Boundary chain"), stream); });
}

int gwn_boundary_clear(gwn_boundary_t boundary) {
    return guarded("gwn_boundary_clear", [&] {
        gwn_boundary_handle &b = *require_handle(boundary, "Boundary chain");
        b.start.release();
        b.end.release();
        b.multiplicity.release();
        b.edge_count = b.mesh_vertex_count = b.mesh_triangle_count = 0;
        b.built = false;
    });
}

int gwn_boundary_get_accessor(gwn_boundary_t boundary, gwn_boundary_accessor *out) {
    return guarded("gwn_boundary_get_accessor", [&] {
        gwn_boundary_handle const &b = *require_handle(boundary, "Boundary chain");
        if (out == nullptr)
            gwn::throw_invalid("Accessor output pointer is null.");
        *out = gwn_boundary_accessor{b.start.ptr, b.end.ptr, b.multiplicity.ptr, b.edge_count,
                                     b.mesh_vertex_count, b.mesh_triangle_count, b.built ? 1 : 0};
    });
}

// ------------------------------------------------------------------------------------------
// moments
// ------------------------------------------------------------------------------------------

// ------------------------------------------------------------------------------------------
// boundary chain
// ------------------------------------------------------------------------------------------

int gwn_boundary_create(gwn_boundary_t *out) {
    return guarded("gwn_boundary_create", [&] {
        if (out == nullptr)
            gwn::throw_invalid("Output handle pointer is null.");
        *out = new gwn_boundary_handle();
    });
}

int gwn_boundary_destroy(gwn_boundary_t boundary) {
    return guarded("gwn_boundary_destroy", [&] { delete boundary; });
}

int gwn_build_boundary_chain(gwn_geometry_t geometry, gwn_boundary_t boundary, hipStream_t stream) {
    return guarded("gwn_build_boundary_chain", [&] {
        if (boundary == nullptr)
            gwn::throw_invalid("Boundary chain handle is null.");
        if (geometry == nullptr || (geometry->triangle_count != 0 && geometry->i0.ptr == nullptr))
            gwn::throw_invalid("Geometry object contains no queryable data.");
        // Rows are placed by a scan with 30-bit partial counts.
        if (geometry->triangle_count > ((1ull << 30) - 1ull) / 3ull)
            gwn::throw_invalid("Boundary chain triangle count exceeds edge encoding.");
        gwn_boundary_handle staging;
        staging.stream = stream;
        gwn::build_boundary_chain(*geometry, staging, stream);
        std::swap(*boundary, staging);
    });
}

int gwn_boundary_info(
    gwn_boundary_t boundary, uint64_t *edge_count, uint64_t *mesh_vertex_count, uint64_t *mesh_triangle_count,
    int32_t *is_built
) {
    return guarded("gwn_boundary_info", [&] {
        if (boundary == nullptr)
            gwn::throw_invalid("Boundary chain handle is null.");
        if (edge_count)
            *edge_count = boundary->edge_count;
        if (mesh_vertex_count)
            *mesh_vertex_count = boundary->mesh_vertex_count;
        if (mesh_triangle_count)
            *mesh_triangle_count = boundary->mesh_triangle_count;
        if (is_built)
            *is_built = boundary->built ? 1 : 0;
    });
}

int gwn_boundary_download(
    gwn_boundary_t boundary, uint32_t *start, uint32_t *end, uint64_t *multiplicity, hipStream_t stream
) {
    return guarded("gwn_boundary_download", [&] {
        if (boundary == nullptr)
            gwn::throw_invalid("Boundary chain handle is null.");
        std::size_t const n = boundary->edge_count;
        if (n == 0)
            return;
        if (start)
            GWN_HIP(hipMemcpyAsync(start, boundary->start.ptr, n * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        if (end)
            GWN_HIP(hipMemcpyAsync(end, boundary->end.ptr, n * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        if (multiplicity)
            GWN_HIP(hipMemcpyAsync(multiplicity, boundary->multiplicity.ptr, n * sizeof(uint64_t),
                                   hipMemcpyDeviceToHost, stream));
        GWN_HIP(hipStreamSynchronize(stream));
    });
}

int gwn_moment_create(gwn_moment_t *out, int32_t order) {
    return guarded("gwn_moment_create", [&] {
        if (out == nullptr)
            gwn::throw_invalid("Output handle pointer is null.");
        if (order < 0 || order > 2)
            gwn::throw_invalid("Taylor order must be 0, 1, or 2.");
        auto *m = new gwn_moment_handle();
        m->order = order;
        *out = m;
    });
}

int gwn_moment_destroy(gwn_moment_t moment) {
    return guarded("gwn_moment_destroy", [&] { delete moment; });
}

int gwn_moment_order(gwn_moment_t moment, int32_t *order) {
    return guarded("gwn_moment_order", [&] {
        if (moment == nullptr || order == nullptr)
            gwn::throw_invalid("Moment handle or output pointer is null.");
        *order = moment->order;
    });
}

int gwn_moment_info(gwn_moment_t moment, uint64_t *node_count, uint64_t *bvh_revision) {
    return guarded("gwn_moment_info", [&] {
        gwn_moment_handle const &m = *require_handle(moment, "Moment");
        if (node_count)
            *node_count = m.node_count;
        if (bvh_revision)
            *bvh_revision = m.bvh_revision;
    });
}

int gwn_refit_bvh_moment_f64(gwn_bvh_t bvh, gwn_moment_t moment, hipStream_t stream) {
    return guarded("gwn_refit_bvh_moment", [&] {
        if (bvh == nullptr || !bvh->valid())
            gwn::throw_invalid("bvh.refit.moment: BVH accessor is invalid for moment refit.");
        if (bvh->real_bytes != 8)
            gwn::throw_invalid("bvh.refit.moment: BVH holds Real = float: use the f32 entry point.");
        if (moment == nullptr)
            gwn::throw_invalid("bvh.refit.moment: Moment handle is null.");
        gwn_moment_handle staging;
        staging.order = moment->order;
        staging.stream = stream;
        staging.storage_generation = gwn::next_storage_generation();
        gwn::refit_moment_double(*bvh, staging, stream);
        std::swap(*moment, staging);
    });
}

int gwn_refit_bvh_moment_f32(gwn_bvh_t bvh, gwn_moment_t moment, hipStream_t stream) {
    return guarded("gwn_refit_bvh_moment", [&] {
        if (bvh == nullptr || !bvh->valid())
            gwn::throw_invalid("bvh.refit.moment: BVH accessor is invalid for moment refit.");
        if (bvh->real_bytes != 4)
            gwn::throw_invalid("bvh.refit.moment: BVH holds Real = double: use the f64 entry point.");
        if (moment == nullptr)
            gwn::throw_invalid("bvh.refit.moment: Moment handle is null.");
        gwn_moment_handle staging;
        staging.order = moment->order;
        staging.width = bvh->width;
        staging.stream = stream;
        staging.storage_generation = gwn::next_storage_generation();
        gwn::refit_moment(*bvh, staging, stream);
        std::swap(*moment, staging);
    });
}

int gwn_moment_download_f64(gwn_moment_t moment, double *fields, hipStream_t stream) {
    return guarded("gwn_moment_download", [&] {
        if (moment == nullptr || fields == nullptr)
            gwn::throw_invalid("Moment handle or output pointer is null.");
        if (moment->real_bytes != 8)
            gwn::throw_invalid("Moment holds Real = float: use the f32 download.");
        std::size_t const N = moment->node_count;
        int const F = gwn::moment_field_count(moment->order);
        if (N == 0 || moment->fields_d.ptr == nullptr)
            return;
        std::vector<double> host(N * F * 4);
        GWN_HIP(hipMemcpyAsync(host.data(), moment->fields_d.ptr, host.size() * sizeof(double),
                               hipMemcpyDeviceToHost, stream));
        GWN_HIP(hipStreamSynchronize(stream));
        // Field-SoA gwn_bvh_moment_node<Width, Order, double>: field f of slot s at [f * Width + s].
        int const W = moment->width;
        for (std::size_t n = 0; n < N; ++n)
            for (int f = 0; f < F; ++f)
                std::memcpy(fields + (n * F + f) * W, host.data() + (n * F + f) * 4, sizeof(double) * W);
    });
}

int gwn_moment_download(gwn_moment_t moment, float *fields, hipStream_t stream) {
    return guarded("gwn_moment_download", [&] {
        if (moment == nullptr || fields == nullptr)
            gwn::throw_invalid("Moment handle or output pointer is null.");
        if (moment->real_bytes != 4)
            gwn::throw_invalid("Moment holds Real = double: use the f64 download.");
        std::size_t const N = moment->node_count;
        int const F = gwn::moment_field_count(moment->order);
        int const RV = gwn::moment_record_vec4(moment->order);
        if (N == 0 || moment->records.ptr == nullptr)
            return;
        std::vector<float> host(N * RV * 4);
        GWN_HIP(hipMemcpyAsync(host.data(), moment->records.ptr, host.size() * sizeof(float), hipMemcpyDeviceToHost, stream));
        GWN_HIP(hipStreamSynchronize(stream));
        // Field-SoA gwn_bvh_moment_node<Width, Order, float>: field f of slot s at [f * Width + s].
        int const W = moment->width;
        for (std::size_t n = 0; n < N; ++n)
            for (int f = 0; f < F; ++f)
                std::memcpy(fields + (n * F + f) * W, host.data() + n * RV * 4 + 4 + f * 4, sizeof(float) * W);
    });
}

// ------------------------------------------------------------------------------------------
// winding-number batches
// ------------------------------------------------------------------------------------------

int gwn_compute_winding_number_exact_batch_f32(
    gwn_bvh_t bvh, const float *qx, const float *qy, const float *qz, float *out, uint64_t count,
    const gwn_query_config *config, hipStream_t stream
) {
    return guarded("gwn_compute_winding_number_exact_batch", [&] {
        gwn_query_config const c = resolve_config(config);
        validate_config(c, false);
        require_bvh(bvh);
        require_query_storage(count, qx, qy, qz, out);
        if (count == 0)
            return;
        gwn::winding_exact(*bvh, qx, qy, qz, out, count, -1, to_launch(c), stream);
    });
}

int gwn_compute_winding_number_exact_batch_ex_f32(
    gwn_bvh_t bvh, const float *qx, const float *qy, const float *qz, double *out, uint64_t count,
    int32_t precision, const gwn_query_config *config, hipStream_t stream
) {
    return guarded("gwn_compute_winding_number_exact_batch", [&] {
        gwn_query_config const c = resolve_config(config);
        validate_config(c, false);
        require_bvh(bvh);
        if (precision != GWN_PRECISION_PLAIN && precision != GWN_PRECISION_SPLIT && precision != GWN_PRECISION_FP64)
            gwn::throw_invalid("Unknown winding precision mode.");
        require_query_storage(count, qx, qy, qz, out);
        if (count == 0)
            return;
        gwn::winding_exact(*bvh, qx, qy, qz, out, count, precision, to_launch(c), stream);
    });
}

// Real = double batches (gwn_query.cuh:140-180, 380-445 instantiated with double): one query per
// lane, the exact sum and the Taylor traversal in double.
int gwn_compute_winding_number_exact_batch_f64(
    gwn_bvh_t bvh, const double *qx, const double *qy, const double *qz, double *out, uint64_t count,
    const gwn_query_config *config, hipStream_t stream
) {
    return guarded("gwn_compute_winding_number_exact_batch", [&] {
        gwn_query_config const c = resolve_config(config);
        validate_config(c, false);
        require_bvh(bvh, 8);
        require_query_storage(count, qx, qy, qz, out);
        if (count == 0)
            return;
        gwn::winding_exact_double(*bvh, qx, qy, qz, out, count, stream);
    });
}

int gwn_compute_winding_number_taylor_batch_f64(
    gwn_bvh_t bvh, gwn_moment_t moment, const double *qx, const double *qy, const double *qz, double *out,
    uint64_t count, double accuracy_scale, const gwn_query_config *config, hipStream_t stream
) {
    return guarded("gwn_compute_winding_number_taylor_batch", [&] {
        gwn_query_config const c = resolve_config(config);
        validate_config(c, true);
        require_bvh(bvh, 8);
        require_stack(*bvh, c.stack_capacity);
        require_moment(moment, *bvh);
        if (!(accuracy_scale >= 0.0))
            gwn::throw_invalid("Taylor accuracy scale must be nonnegative.");
        require_query_storage(count, qx, qy, qz, out);
        if (count == 0)
            return;
        gwn::winding_taylor_double(*bvh, *moment, qx, qy, qz, out, count, accuracy_scale, stream);
    });
}

// Shared validation of the Taylor winding and gradient batches, in the reference's order
// (gwn_query.cuh:409-428, 711-737).  `out` may be null only when the gradient is requested.
static void taylor_common(
    gwn_bvh_t bvh, gwn_moment_t moment, const float *qx, const float *qy, const float *qz, float *out,
    float *gx, float *gy, float *gz, uint64_t count, float accuracy_scale, int32_t precision,
    const gwn_query_config *config, const uint32_t *order, uint32_t *counters, hipStream_t stream
) {
    gwn_query_config const c = resolve_config(config);
    validate_config(c, true);
    require_bvh(bvh);
    require_stack(*bvh, c.stack_capacity);
    require_moment(moment, *bvh);
    if (!(accuracy_scale >= 0.0f))
        gwn::throw_invalid("Taylor accuracy scale must be nonnegative.");
    if (precision != GWN_PRECISION_PLAIN && precision != GWN_PRECISION_SPLIT)
        gwn::throw_invalid("Taylor winding supports the plain and split precision modes.");
    if (gx != nullptr || gy != nullptr || gz != nullptr) {
        require_query_storage(count, qx, qy, qz, gx);
        if (count != 0 && (gy == nullptr || gz == nullptr))
            gwn::throw_invalid("Query/output spans must use non-null storage when non-empty.");
    } else {
        require_query_storage(count, qx, qy, qz, out);
    }
    if (count == 0)
        return;
    gwn::winding_taylor(*bvh, *moment, qx, qy, qz, out, gx, gy, gz, count, accuracy_scale, precision, to_launch(c),
                        order, counters, stream);
}

int gwn_compute_winding_number_taylor_batch_f32(
    gwn_bvh_t bvh, gwn_moment_t moment, const float *qx, const float *qy, const float *qz, float *out,
    uint64_t count, float accuracy_scale, const gwn_query_config *config, hipStream_t stream
) {
    return guarded("gwn_compute_winding_number_taylor_batch", [&] {
        taylor_common(bvh, moment, qx, qy, qz, out, nullptr, nullptr, nullptr, count, accuracy_scale,
                      GWN_PRECISION_PLAIN, config, nullptr, nullptr, stream);
    });
}

int gwn_compute_winding_number_taylor_batch_ex_f32(
    gwn_bvh_t bvh, gwn_moment_t moment, const float *qx, const float *qy, const float *qz, float *out,
    uint64_t count, float accuracy_scale, int32_t precision, const gwn_query_config *config,
    const uint32_t *order, uint32_t *counters, hipStream_t stream
) {
    return guarded("gwn_compute_winding_number_taylor_batch", [&] {
        taylor_common(bvh, moment, qx, qy, qz, out, nullptr, nullptr, nullptr, count, accuracy_scale, precision,
                      config, order, counters, stream);
    });
}

int gwn_compute_winding_number_taylor_batch_noalloc_f32(
    gwn_bvh_t bvh, gwn_moment_t moment, const float *qx, const float *qy, const float *qz, float *out,
    uint64_t count, float accuracy_scale, int32_t precision, const gwn_query_config *config, hipStream_t stream
) {
    return guarded("gwn_compute_winding_number_taylor_batch_noalloc", [&] {
        gwn_query_config const c = resolve_config(config);
        validate_config(c, true);
        require_bvh(bvh);
        require_stack(*bvh, c.stack_capacity);
        require_moment(moment, *bvh);
        if (!(accuracy_scale >= 0.0f))
            gwn::throw_invalid("Taylor accuracy scale must be nonnegative.");
        if (precision != GWN_PRECISION_PLAIN && precision != GWN_PRECISION_SPLIT)
            gwn::throw_invalid("Taylor winding supports the plain and split precision modes.");
        require_query_storage(count, qx, qy, qz, out);
        if (count > gwn::k_noalloc_max_queries)
            gwn::throw_invalid("noalloc Taylor batch: query count exceeds the static scratch capacity (1048576 queries).");
        gwn::winding_taylor_noalloc(*bvh, *moment, qx, qy, qz, out, count, accuracy_scale, precision, to_launch(c),
                                    stream);
    });
}

int gwn_compute_winding_gradient_taylor_batch_f32(
    gwn_bvh_t bvh, gwn_moment_t moment, const float *qx, const float *qy, const float *qz, float *gx,
    float *gy, float *gz, uint64_t count, float accuracy_scale, const gwn_query_config *config,
    hipStream_t stream
) {
    return guarded("gwn_compute_winding_gradient_taylor_batch", [&] {
        if (gx == nullptr && gy == nullptr && gz == nullptr && count != 0)
            gwn::throw_invalid("Query/output spans must use non-null storage when non-empty.");
        taylor_common(bvh, moment, qx, qy, qz, nullptr, gx, gy, gz, count, accuracy_scale, GWN_PRECISION_PLAIN,
                      config, nullptr, nullptr, stream);
    });
}

int gwn_compute_winding_gradient_taylor_batch_ex_f32(
    gwn_bvh_t bvh, gwn_moment_t moment, const float *qx, const float *qy, const float *qz, float *out,
    float *gx, float *gy, float *gz, uint64_t count, float accuracy_scale, int32_t precision,
    const gwn_query_config *config, const uint32_t *order, uint32_t *counters, hipStream_t stream
) {
    return guarded("gwn_compute_winding_gradient_taylor_batch", [&] {
        if (gx == nullptr && gy == nullptr && gz == nullptr && count != 0)
            gwn::throw_invalid("Query/output spans must use non-null storage when non-empty.");
        taylor_common(bvh, moment, qx, qy, qz, out, gx, gy, gz, count, accuracy_scale, precision, config, order,
                      counters, stream);
    });
}

int gwn_compute_unsigned_distance_batch_f32(
    gwn_bvh_t bvh, const float *qx, const float *qy, const float *qz, float *out, uint64_t count,
    float culling_band, const gwn_query_config *config, hipStream_t stream
) {
    return guarded("gwn_compute_unsigned_distance_batch", [&] {
        gwn_query_config const c = resolve_config(config);
        validate_config(c, true);
        validate_lanes(c);
        require_bvh(bvh);
        require_stack(*bvh, c.stack_capacity);
        require_query_storage(count, qx, qy, qz, out);
        if (count == 0)
            return;
        gwn::unsigned_distance(*bvh, qx, qy, qz, out, count, culling_band, to_launch(c), nullptr, stream);
    });
}

int gwn_compute_unsigned_distance_batch_ex_f32(
    gwn_bvh_t bvh, const float *qx, const float *qy, const float *qz, float *out, uint32_t *nearest, uint64_t count,
    float culling_band, const gwn_query_config *config, hipStream_t stream
) {
    return guarded("gwn_compute_unsigned_distance_batch", [&] {
        gwn_query_config const c = resolve_config(config);
        validate_config(c, true);
        validate_lanes(c);
        require_bvh(bvh);
        require_stack(*bvh, c.stack_capacity);
        require_query_storage(count, qx, qy, qz, out);
        if (count != 0 && nearest == nullptr)
            gwn::throw_invalid("Nearest-record output is null.");
        if (count == 0)
            return;
        gwn::unsigned_distance(*bvh, qx, qy, qz, out, count, culling_band, to_launch(c), nearest, stream);
    });
}

int gwn_compute_ray_first_hit_batch_f32(
    gwn_bvh_t bvh, const float *ox, const float *oy, const float *oz, const float *dx, const float *dy,
    const float *dz, gwn_ray_first_hit_result *out, uint64_t count, float t_min, float t_max,
    const gwn_query_config *config, hipStream_t stream
) {
    return guarded("gwn_compute_ray_first_hit_batch", [&] {
        gwn_query_config const c = resolve_config(config);
        validate_config(c, true);
        require_bvh(bvh);
        require_stack(*bvh, c.stack_capacity);
        if (count != 0 && (ox == nullptr || oy == nullptr || oz == nullptr || dx == nullptr || dy == nullptr ||
                           dz == nullptr || out == nullptr))
            gwn::throw_invalid("ray first-hit: ray/output spans must use non-null storage when non-empty.");
        if (!(t_max >= t_min))
            gwn::throw_invalid("ray first-hit: requires t_max >= t_min.");
        if (count == 0)
            return;
        gwn::ray_first_hit(*bvh, ox, oy, oz, dx, dy, dz, out, count, t_min, t_max, to_launch(c), stream);
    });
}

int gwn_compute_unsigned_distance_work_f32(
    gwn_bvh_t bvh, const float *qx, const float *qy, const float *qz, uint64_t count, float culling_band,
    uint32_t *counters, hipStream_t stream
) {
    return guarded("gwn_compute_unsigned_distance_work", [&] {
        require_bvh(bvh);
        require_query_storage(count, qx, qy, qz, counters);
        if (count == 0)
            return;
        gwn::unsigned_distance_work(*bvh, qx, qy, qz, count, culling_band, counters, stream);
    });
}

int gwn_compute_ray_first_hit_work_f32(
    gwn_bvh_t bvh, const float *ox, const float *oy, const float *oz, const float *dx, const float *dy,
    const float *dz, uint64_t count, float t_min, float t_max, const gwn_query_config *config, uint32_t *counters,
    hipStream_t stream
) {
    return guarded("gwn_compute_ray_first_hit_work", [&] {
        gwn_query_config const c = resolve_config(config);
        validate_config(c, true);
        require_bvh(bvh);
        require_stack(*bvh, c.stack_capacity);
        if (count != 0 && (ox == nullptr || oy == nullptr || oz == nullptr || dx == nullptr || dy == nullptr ||
                           dz == nullptr || counters == nullptr))
            gwn::throw_invalid("ray first-hit: ray/output spans must use non-null storage when non-empty.");
        if (!(t_max >= t_min))
            gwn::throw_invalid("ray first-hit: requires t_max >= t_min.");
        if (count == 0)
            return;
        gwn::ray_first_hit_work(*bvh, ox, oy, oz, dx, dy, dz, count, t_min, t_max, to_launch(c), counters, stream);
    });
}

// Shared validation of the antipodal batches, in the reference's order (gwn_query.cuh:510-540,
// 606-630).  `bvh` is null for the gradient, which reads the boundary term only.
static void require_antipodal(
    gwn_geometry_t geometry, gwn_bvh_t bvh, bool needs_bvh, gwn_boundary_t boundary, int32_t precision,
    gwn_query_config const &c
) {
    if (geometry == nullptr || (geometry->triangle_count != 0 && geometry->i0.ptr == nullptr))
        gwn::throw_invalid("Geometry object contains no queryable data.");
    require_geometry_real(geometry, 4);
    if (needs_bvh) {
        require_bvh(bvh);
        if (bvh->triangle_count != geometry->triangle_count)
            gwn::throw_invalid("Geometry and BVH triangle counts differ.");
        require_stack(*bvh, c.stack_capacity);
    }
    if (boundary == nullptr || !boundary->built)
        gwn::throw_invalid("Boundary chain object is unbuilt or invalid.");
    if (boundary->mesh_vertex_count != geometry->vertex_count ||
        boundary->mesh_triangle_count != geometry->triangle_count)
        gwn::throw_invalid("Boundary chain belongs to a different mesh shape.");
    if (precision != GWN_PRECISION_PLAIN && precision != GWN_PRECISION_SPLIT)
        gwn::throw_invalid("Antipodal queries support the plain and split precision modes.");
}

int gwn_compute_winding_number_antipodal_batch_ex_f32(
    gwn_geometry_t geometry, gwn_bvh_t bvh, gwn_boundary_t boundary, const float *qx, const float *qy,
    const float *qz, float *out, uint64_t count, int32_t precision, const gwn_query_config *config,
    hipStream_t stream
) {
    return guarded("gwn_compute_winding_number_antipodal_batch", [&] {
        gwn_query_config const c = resolve_config(config);
        validate_config(c, true);
        require_antipodal(geometry, bvh, true, boundary, precision, c);
        require_query_storage(count, qx, qy, qz, out);
        if (count == 0)
            return;
        gwn::winding_antipodal(*geometry, *bvh, *boundary, qx, qy, qz, out, count, precision, stream);
    });
}

int gwn_compute_winding_number_antipodal_batch_f32(
    gwn_geometry_t geometry, gwn_bvh_t bvh, gwn_boundary_t boundary, const float *qx, const float *qy,
    const float *qz, float *out, uint64_t count, const gwn_query_config *config, hipStream_t stream
) {
    return gwn_compute_winding_number_antipodal_batch_ex_f32(geometry, bvh, boundary, qx, qy, qz, out, count,
                                                             GWN_PRECISION_PLAIN, config, stream);
}

int gwn_compute_winding_gradient_antipodal_batch_ex_f32(
    gwn_geometry_t geometry, gwn_boundary_t boundary, const float *qx, const float *qy, const float *qz,
    float *gx, float *gy, float *gz, uint64_t count, int32_t precision, const gwn_query_config *config,
    hipStream_t stream
) {
    return guarded("gwn_compute_winding_gradient_antipodal_batch", [&] {
        gwn_query_config const c = resolve_config(config);
        validate_config(c, false);
        require_antipodal(geometry, nullptr, false, boundary, precision, c);
        require_query_storage(count, qx, qy, qz, gx);
        if (count != 0 && (gy == nullptr || gz == nullptr))
            gwn::throw_invalid("Query/output spans must use non-null storage when non-empty.");
        if (count == 0)
            return;
        gwn::winding_gradient_antipodal(*geometry, *boundary, qx, qy, qz, gx, gy, gz, count, precision, stream);
    });
}

int gwn_compute_winding_gradient_antipodal_batch_f32(
    gwn_geometry_t geometry, gwn_boundary_t boundary, const float *qx, const float *qy, const float *qz,
    float *gx, float *gy, float *gz, uint64_t count, const gwn_query_config *config, hipStream_t stream
) {
    return gwn_compute_winding_gradient_antipodal_batch_ex_f32(geometry, boundary, qx, qy, qz, gx, gy, gz, count,
                                                               GWN_PRECISION_PLAIN, config, stream);
}

// Adjoint batches (no reference counterpart).  Validation follows the forward batches they
// differentiate; the vertex gradient arrays are added to, so several losses can share them.
static void require_adjoint_geometry(gwn_geometry_t geometry, gwn_bvh_t bvh, gwn_query_config const &c,
                                     float const *gvx, float const *gvy, float const *gvz) {
    if (geometry == nullptr || (geometry->triangle_count != 0 && geometry->i0.ptr == nullptr))
        gwn::throw_invalid("Geometry object contains no queryable data.");
    require_geometry_real(geometry, 4);
    require_bvh(bvh);
    if (bvh->triangle_count != geometry->triangle_count)
        gwn::throw_invalid("Geometry and BVH triangle counts differ.");
    require_stack(*bvh, c.stack_capacity);
    if (geometry->vertex_count != 0 && (gvx == nullptr || gvy == nullptr || gvz == nullptr))
        gwn::throw_invalid("Vertex gradient spans must use non-null storage.");
}

int gwn_compute_winding_taylor_adjoint_batch_f32(
    gwn_geometry_t geometry, gwn_bvh_t bvh, gwn_moment_t moment, const float *qx, const float *qy,
    const float *qz, const float *grad_w, uint64_t count, float accuracy_scale, float *grad_vx, float *grad_vy,
    float *grad_vz, float *grad_qx, float *grad_qy, float *grad_qz, const gwn_query_config *config,
    hipStream_t stream
) {
    return guarded("gwn_compute_winding_taylor_adjoint_batch", [&] {
        gwn_query_config const c = resolve_config(config);
        validate_config(c, true);
        require_adjoint_geometry(geometry, bvh, c, grad_vx, grad_vy, grad_vz);
        require_moment(moment, *bvh);
        if (moment->order > 1)
            gwn::throw_invalid("The winding adjoint supports Taylor orders 0 and 1.");
        if (!(accuracy_scale >= 0.0f))
            gwn::throw_invalid("Taylor accuracy scale must be nonnegative.");
        require_query_storage(count, qx, qy, qz, grad_w);
        if ((grad_qx != nullptr || grad_qy != nullptr || grad_qz != nullptr) && count != 0 &&
            (grad_qx == nullptr || grad_qy == nullptr || grad_qz == nullptr))
            gwn::throw_invalid("Query gradient spans must be all null or all non-null.");
        if (count == 0)
            return;
        gwn::winding_taylor_adjoint(*geometry, *bvh, *moment, qx, qy, qz, grad_w, count, accuracy_scale, grad_vx,
                                    grad_vy, grad_vz, grad_qx, grad_qy, grad_qz, stream);
    });
}

int gwn_compute_unsigned_distance_adjoint_batch_f32(
    gwn_geometry_t geometry, gwn_bvh_t bvh, const float *px, const float *py, const float *pz,
    const float *grad_d, uint64_t count, float culling_band, float *grad_vx, float *grad_vy, float *grad_vz,
    float *grad_px, float *grad_py, float *grad_pz, const gwn_query_config *config, hipStream_t stream
) {
    return guarded("gwn_compute_unsigned_distance_adjoint_batch", [&] {
        gwn_query_config const c = resolve_config(config);
        validate_config(c, true);
        require_adjoint_geometry(geometry, bvh, c, grad_vx, grad_vy, grad_vz);
        require_query_storage(count, px, py, pz, grad_d);
        if ((grad_px != nullptr || grad_py != nullptr || grad_pz != nullptr) && count != 0 &&
            (grad_px == nullptr || grad_py == nullptr || grad_pz == nullptr))
            gwn::throw_invalid("Point gradient spans must be all null or all non-null.");
        if (count == 0)
            return;
        gwn::unsigned_distance_adjoint(*geometry, *bvh, px, py, pz, grad_d, count, culling_band, nullptr, grad_vx,
                                       grad_vy, grad_vz, grad_px, grad_py, grad_pz, stream);
    });
}

int gwn_compute_unsigned_distance_adjoint_batch_ex_f32(
    gwn_geometry_t geometry, gwn_bvh_t bvh, const float *px, const float *py, const float *pz,
    const uint32_t *nearest, const float *grad_d, uint64_t count, float *grad_vx, float *grad_vy, float *grad_vz,
    float *grad_px, float *grad_py, float *grad_pz, hipStream_t stream
) {
    return guarded("gwn_compute_unsigned_distance_adjoint_batch", [&] {
        gwn_query_config const c = resolve_config(nullptr);
        require_adjoint_geometry(geometry, bvh, c, grad_vx, grad_vy, grad_vz);
        require_query_storage(count, px, py, pz, grad_d);
        if (count != 0 && nearest == nullptr)
            gwn::throw_invalid("Nearest-record input is null.");
        if ((grad_px != nullptr || grad_py != nullptr || grad_pz != nullptr) && count != 0 &&
            (grad_px == nullptr || grad_py == nullptr || grad_pz == nullptr))
            gwn::throw_invalid("Point gradient spans must be all null or all non-null.");
        if (count == 0)
            return;
        gwn::unsigned_distance_adjoint(*geometry, *bvh, px, py, pz, grad_d, count, 0.0f, nearest, grad_vx, grad_vy,
                                       grad_vz, grad_px, grad_py, grad_pz, stream);
    });
}

// ------------------------------------------------------------------------------------------
// batched mode (no reference counterpart)
// ------------------------------------------------------------------------------------------

int gwn_batch_create(const gwn_batch_item *items, uint64_t item_count, float accuracy_scale, gwn_batch_t *out,
                     hipStream_t stream) {
    return guarded("gwn_batch_create", [&] {
        if (out == nullptr)
            gwn::throw_invalid("Output handle pointer is null.");
        if (item_count != 0 && items == nullptr)
            gwn::throw_invalid("Batch item array is null.");
        if (item_count > 0xffffffffull)
            gwn::throw_invalid("Batch item count exceeds the mesh index range.");
        if (!(accuracy_scale >= 0.0f))
            gwn::throw_invalid("Taylor accuracy scale must be nonnegative.");
        std::vector<gwn_bvh_handle const *> bvhs;
        std::vector<gwn_moment_handle *> moments;
        std::vector<gwn::batch_mesh> qs;
        for (uint64_t i = 0; i < item_count; ++i) {
            gwn_batch_item const &it = items[i];
            require_bvh(it.bvh);
            if (it.moment == nullptr)
                gwn::throw_invalid("Moment object is not aligned with the BVH.");
            if (i > 0 && it.moment->order != items[0].moment->order)
                gwn::throw_invalid("batch: Every moment object must have the same Taylor order.");
            require_query_storage(it.count, it.qx, it.qy, it.qz, it.out);
            bvhs.push_back(it.bvh);
            moments.push_back(it.moment);
            gwn::batch_mesh m{};
            m.qx = it.qx;
            m.qy = it.qy;
            m.qz = it.qz;
            m.out = it.out;
            m.count = it.count;
            qs.push_back(m);
        }
        *out = gwn::batch_create(bvhs, moments, qs, accuracy_scale, stream);
    });
}

int gwn_batch_destroy(gwn_batch_t batch) {
    return guarded("gwn_batch_destroy", [&] { gwn::batch_destroy(batch); });
}

int gwn_batch_refit_moments(gwn_batch_t batch, hipStream_t stream) {
    return guarded("gwn_batch_refit_moments", [&] {
        if (batch == nullptr)
            gwn::throw_invalid("Batch handle is null.");
        gwn::batch_refit(*batch, stream);
    });
}

int gwn_batch_query(gwn_batch_t batch, hipStream_t stream) {
    return guarded("gwn_batch_query", [&] {
        if (batch == nullptr)
            gwn::throw_invalid("Batch handle is null.");
        gwn::batch_query(*batch, stream);
    });
}

int gwn_batch_set_tuning(gwn_batch_t batch, const gwn_batch_tuning *tuning, hipStream_t stream) {
    return guarded("gwn_batch_set_tuning", [&] {
        if (batch == nullptr || tuning == nullptr)
            gwn::throw_invalid("Batch handle or tuning pointer is null.");
        gwn::batch_set_tuning(*batch, tuning->block_size, tuning->chunk, tuning->blocks_per_cu, tuning->traversal,
                              tuning->wave_queries_per_triangle, tuning->stack_capacity, stream);
    });
}

int gwn_batch_get_tuning(gwn_batch_t batch, gwn_batch_tuning *tuning) {
    return guarded("gwn_batch_get_tuning", [&] {
        if (batch == nullptr || tuning == nullptr)
            gwn::throw_invalid("Batch handle or tuning pointer is null.");
        gwn::batch_get_tuning(*batch, *tuning);
    });
}

int gwn_batch_tune(gwn_batch_t batch, hipStream_t stream, float *best_ms) {
    return guarded("gwn_batch_tune", [&] {
        if (batch == nullptr)
            gwn::throw_invalid("Batch handle is null.");
        float const ms = gwn::batch_tune(*batch, stream);
        if (best_ms)
            *best_ms = ms;
    });
}

int gwn_batch_capture(gwn_batch_t batch, int32_t include_refit) {
    return guarded("gwn_batch_capture", [&] {
        if (batch == nullptr)
            gwn::throw_invalid("Batch handle is null.");
        gwn::batch_capture(*batch, include_refit != 0);
    });
}

int gwn_batch_replay(gwn_batch_t batch, hipStream_t stream) {
    return guarded("gwn_batch_replay", [&] {
        if (batch == nullptr)
            gwn::throw_invalid("Batch handle is null.");
        gwn::batch_replay(*batch, stream);
    });
}

int gwn_batch_status(gwn_batch_t batch, hipStream_t stream) {
    return guarded("gwn_batch_status", [&] {
        if (batch == nullptr)
            gwn::throw_invalid("Batch handle is null.");
        std::uint32_t const bits = gwn::batch_status(*batch, stream);
        if (bits != 0u)
            gwn::throw_internal(gwn::batch_status_message(bits));
    });
}

int gwn_batch_create_built(const gwn_batch_build_item *items, uint64_t item_count, const gwn_batch_build_config *config,
                           float accuracy_scale, gwn_batch_t *out, hipStream_t stream) {
    return guarded("gwn_batch_create_built", [&] {
        if (out == nullptr)
            gwn::throw_invalid("Output handle pointer is null.");
        if (config == nullptr)
            gwn::throw_invalid("Batch build configuration pointer is null.");
        if (item_count != 0 && items == nullptr)
            gwn::throw_invalid("Batch item array is null.");
        if (item_count > 0xffffffffull)
            gwn::throw_invalid("Batch item count exceeds the mesh index range.");
        if (!(accuracy_scale >= 0.0f))
            gwn::throw_invalid("Taylor accuracy scale must be nonnegative.");
        std::vector<gwn_geometry_handle const *> geometries;
        std::vector<gwn::batch_mesh> qs;
        for (uint64_t i = 0; i < item_count; ++i) {
            gwn_batch_build_item const &it = items[i];
            if (it.geometry == nullptr || (it.geometry->triangle_count != 0 && it.geometry->i0.ptr == nullptr))
                gwn::throw_invalid("bvh.build: Geometry accessor is invalid for BVH build.");
            require_geometry_real(it.geometry, 4);
            require_query_storage(it.count, it.qx, it.qy, it.qz, it.out);
            geometries.push_back(it.geometry);
            gwn::batch_mesh m{};
            m.qx = it.qx;
            m.qy = it.qy;
            m.qz = it.qz;
            m.out = it.out;
            m.count = it.count;
            qs.push_back(m);
        }
        *out = gwn::batch_create_built(geometries, qs, config->method, config->width, config->hploc_search_radius,
                                       config->order, accuracy_scale, stream);
    });
}

int gwn_batch_build(gwn_batch_t batch, hipStream_t stream) {
    return guarded("gwn_batch_build", [&] {
        if (batch == nullptr)
            gwn::throw_invalid("Batch handle is null.");
        gwn::batch_build(*batch, stream);
    });
}

int gwn_batch_set_build(gwn_batch_t batch, const gwn_batch_build_config *config, hipStream_t stream) {
    return guarded("gwn_batch_set_build", [&] {
        if (batch == nullptr || config == nullptr)
            gwn::throw_invalid("Batch handle or build configuration pointer is null.");
        gwn::batch_set_build(*batch, config->method, config->width, config->hploc_search_radius, stream);
    });
}

int gwn_batch_get_build(gwn_batch_t batch, gwn_batch_build_config *config, int32_t *level_launches) {
    return guarded("gwn_batch_get_build", [&] {
        if (batch == nullptr || config == nullptr)
            gwn::throw_invalid("Batch handle or build configuration pointer is null.");
        int method = 0, width = 0, levels = 0, order = 0;
        std::uint32_t radius = 0;
        gwn::batch_get_build(*batch, method, width, radius, levels, order);
        config->method = method;
        config->order = order;
        config->width = width;
        config->hploc_search_radius = radius;
        if (level_launches)
            *level_launches = levels;
    });
}

int gwn_batch_tune_build(gwn_batch_t batch, hipStream_t stream, float *best_ms) {
    return guarded("gwn_batch_tune_build", [&] {
        if (batch == nullptr)
            gwn::throw_invalid("Batch handle is null.");
        float const ms = gwn::batch_tune_build(*batch, stream);
        if (best_ms)
            *best_ms = ms;
    });
}

int gwn_batch_export(gwn_batch_t batch, uint64_t mesh, gwn_bvh_t bvh, gwn_moment_t moment, hipStream_t stream) {
    return guarded("gwn_batch_export", [&] {
        if (batch == nullptr || bvh == nullptr || moment == nullptr)
            gwn::throw_invalid("Batch, BVH or moment handle is null.");
        gwn::batch_export(*batch, mesh, *bvh, *moment, stream);
    });
}

int gwn_compute_query_order_f32(
    gwn_bvh_t bvh, const float *qx, const float *qy, const float *qz, uint32_t *order, uint64_t count,
    hipStream_t stream
) {
    return guarded("gwn_compute_query_order", [&] {
        require_bvh(bvh);
        require_query_storage(count, qx, qy, qz, order);
        if (count > 0x3fffffffull)
            gwn::throw_invalid("Query count exceeds the ordering range (2^30 - 1).");
        gwn::query_order(*bvh, qx, qy, qz, order, count, stream);
    });
}

// ------------------------------------------------------------------------------------------
// sort building block
// ------------------------------------------------------------------------------------------

int gwn_radix_sort_pairs_u64(uint64_t *keys, uint32_t *values, uint64_t count, int32_t begin_bit, int32_t end_bit, hipStream_t stream) {
    return guarded("gwn_radix_sort_pairs_u64", [&] {
        if (count != 0 && (keys == nullptr || values == nullptr))
            gwn::throw_invalid("Sort spans must use non-null storage when non-empty.");
        if (begin_bit < 0 || end_bit > 64 || begin_bit > end_bit)
            gwn::throw_invalid("Sort bit range must satisfy 0 <= begin_bit <= end_bit <= 64.");
        gwn::device_buffer<std::uint64_t> kalt(count, stream);
        gwn::device_buffer<std::uint32_t> valt(count, stream);
        std::uint64_t *ko = nullptr;
        std::uint32_t *vo = nullptr;
        gwn::radix_sort_pairs_u64(keys, kalt.ptr, values, valt.ptr, count, begin_bit, end_bit, &ko, &vo, stream);
        if (ko != keys) {
            GWN_HIP(hipMemcpyAsync(keys, ko, count * sizeof(std::uint64_t), hipMemcpyDeviceToDevice, stream));
            GWN_HIP(hipMemcpyAsync(values, vo, count * sizeof(std::uint32_t), hipMemcpyDeviceToDevice, stream));
        }
    });
}

int gwn_radix_sort_pairs_u32(uint32_t *keys, uint32_t *values, uint64_t count, int32_t begin_bit, int32_t end_bit, hipStream_t stream) {
    return guarded("gwn_radix_sort_pairs_u32", [&] {
        if (count != 0 && (keys == nullptr || values == nullptr))
            gwn::throw_invalid("Sort spans must use non-null storage when non-empty.");
        if (begin_bit < 0 || end_bit > 32 || begin_bit > end_bit)
            gwn::throw_invalid("Sort bit range must satisfy 0 <= begin_bit <= end_bit <= 32.");
        gwn::device_buffer<std::uint32_t> kalt(count, stream);
        gwn::device_buffer<std::uint32_t> valt(count, stream);
        std::uint32_t *ko = nullptr;
        std::uint32_t *vo = nullptr;
        gwn::radix_sort_pairs_u32(keys, kalt.ptr, values, valt.ptr, count, begin_bit, end_bit, &ko, &vo, stream);
        if (ko != keys) {
            GWN_HIP(hipMemcpyAsync(keys, ko, count * sizeof(std::uint32_t), hipMemcpyDeviceToDevice, stream));
            GWN_HIP(hipMemcpyAsync(values, vo, count * sizeof(std::uint32_t), hipMemcpyDeviceToDevice, stream));
        }
    });
}

} // extern "C"
