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

void require_bvh(gwn_bvh_handle const *bvh) {
    if (bvh == nullptr || !bvh->valid())
        gwn::throw_invalid("BVH object contains no queryable data.");
}

void require_stack(gwn_bvh_handle const &bvh, int capacity) {
    if (gwn::ref_internal(bvh.root.reference) && bvh.internal_stack_bound > static_cast<std::uint32_t>(capacity))
        gwn::throw_invalid("Traversal stack capacity is below the BVH internal stack bound.");
}

void require_moment(gwn_moment_handle const *m, gwn_bvh_handle const &bvh) {
    bool ok = m != nullptr && m->bvh_revision != 0 && m->bvh_revision == bvh.revision;
    if (ok) {
        if (gwn::ref_leaf(bvh.root.reference))
            ok = m->node_count == 0;
        else
            ok = m->node_count == bvh.node_count && m->records.ptr != nullptr;
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

int gwn_upload_geometry_f32(
    gwn_geometry_t geometry, const float *x, const float *y, const float *z, uint64_t vertex_count,
    const uint32_t *i0, const uint32_t *i1, const uint32_t *i2, uint64_t triangle_count, hipStream_t stream
) {
    return guarded("gwn_upload_geometry", [&] {
        if (geometry == nullptr)
            gwn::throw_invalid("Geometry handle is null.");
        if (vertex_count > 0xffffffffull)
            gwn::throw_invalid("Vertex count exceeds index type capacity.");
        if (triangle_count > 0xffffffffull)
            gwn::throw_invalid("Triangle count exceeds index type capacity.");
        if ((vertex_count != 0 && (x == nullptr || y == nullptr || z == nullptr)) ||
            (triangle_count != 0 && (i0 == nullptr || i1 == nullptr || i2 == nullptr)))
            gwn::throw_invalid("Geometry spans must use non-null storage when non-empty.");
        for (std::uint64_t t = 0; t < triangle_count; ++t)
            if (i0[t] >= vertex_count || i1[t] >= vertex_count || i2[t] >= vertex_count)
                gwn::throw_invalid("Triangle indices must be in [0, vertex_count).");
        gwn_geometry_handle staging;
        staging.stream = stream;
        staging.vertex_count = vertex_count;
        staging.triangle_count = triangle_count;
        staging.x.allocate(vertex_count, stream);
        staging.y.allocate(vertex_count, stream);
        staging.z.allocate(vertex_count, stream);
        staging.i0.allocate(triangle_count, stream);
        staging.i1.allocate(triangle_count, stream);
        staging.i2.allocate(triangle_count, stream);
        std::size_t const vb = vertex_count * sizeof(float), tb = triangle_count * sizeof(std::uint32_t);
        if (vb) {
            GWN_HIP(hipMemcpyAsync(staging.x.ptr, x, vb, hipMemcpyHostToDevice, stream));
            GWN_HIP(hipMemcpyAsync(staging.y.ptr, y, vb, hipMemcpyHostToDevice, stream));
            GWN_HIP(hipMemcpyAsync(staging.z.ptr, z, vb, hipMemcpyHostToDevice, stream));
        }
        if (tb) {
            GWN_HIP(hipMemcpyAsync(staging.i0.ptr, i0, tb, hipMemcpyHostToDevice, stream));
            GWN_HIP(hipMemcpyAsync(staging.i1.ptr, i1, tb, hipMemcpyHostToDevice, stream));
            GWN_HIP(hipMemcpyAsync(staging.i2.ptr, i2, tb, hipMemcpyHostToDevice, stream));
        }
        // Host sources may be pageable and released by the caller after return.
        GWN_HIP(hipStreamSynchronize(stream));
        staging.generation = geometry->generation + 1;
        std::swap(staging.x, geometry->x);
        std::swap(staging.y, geometry->y);
        std::swap(staging.z, geometry->z);
        std::swap(staging.i0, geometry->i0);
        std::swap(staging.i1, geometry->i1);
        std::swap(staging.i2, geometry->i2);
        geometry->vertex_count = staging.vertex_count;
        geometry->triangle_count = staging.triangle_count;
        geometry->stream = stream;
        geometry->generation = staging.generation;
    });
}

static void update_positions(
    gwn_geometry_t geometry, const float *x, const float *y, const float *z, uint64_t vertex_count,
    hipStream_t stream, hipMemcpyKind kind
) {
    if (geometry == nullptr || (geometry->vertex_count != 0 && geometry->x.ptr == nullptr))
        gwn::throw_invalid("Geometry accessor is invalid for update.");
    if (vertex_count != geometry->vertex_count)
        gwn::throw_invalid("Updated vertex count must match geometry.");
    if (vertex_count != 0 && (x == nullptr || y == nullptr || z == nullptr))
        gwn::throw_invalid("Geometry update spans must use non-null storage when non-empty.");
    geometry->stream = stream;
    std::size_t const vb = vertex_count * sizeof(float);
    if (vb) {
        GWN_HIP(hipMemcpyAsync(geometry->x.ptr, x, vb, kind, stream));
        GWN_HIP(hipMemcpyAsync(geometry->y.ptr, y, vb, kind, stream));
        GWN_HIP(hipMemcpyAsync(geometry->z.ptr, z, vb, kind, stream));
    }
    geometry->generation += 1;
}

int gwn_update_geometry_f32(
    gwn_geometry_t geometry, const float *x, const float *y, const float *z, uint64_t vertex_count, hipStream_t stream
) {
    return guarded("gwn_update_geometry", [&] {
        update_positions(geometry, x, y, z, vertex_count, stream, hipMemcpyHostToDevice);
        GWN_HIP(hipStreamSynchronize(stream));
    });
}

int gwn_update_geometry_device_f32(
    gwn_geometry_t geometry, const float *x, const float *y, const float *z, uint64_t vertex_count, hipStream_t stream
) {
    return guarded("gwn_update_geometry", [&] {
        update_positions(geometry, x, y, z, vertex_count, stream, hipMemcpyDeviceToDevice);
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

int gwn_build_bvh4_f32(
    gwn_geometry_t geometry, gwn_bvh_t bvh, int32_t method, uint32_t hploc_search_radius, hipStream_t stream
) {
    return guarded("gwn_build_bvh", [&] {
        if (bvh == nullptr)
            gwn::throw_invalid("bvh.build: BVH handle is null.");
        if (geometry == nullptr || (geometry->triangle_count != 0 && geometry->i0.ptr == nullptr))
            gwn::throw_invalid("bvh.build: Geometry accessor is invalid for BVH build.");
        if (method != GWN_BVH_BUILD_HPLOC && method != GWN_BVH_BUILD_LBVH)
            gwn::throw_invalid("bvh.build: BVH build method is invalid.");
        if (method == GWN_BVH_BUILD_HPLOC && (hploc_search_radius == 0 || hploc_search_radius > 8))
            gwn::throw_invalid("bvh.build: H-PLOC search radius must be in the range [1, 8].");
        if (geometry->triangle_count != 0 && geometry->vertex_count == 0)
            gwn::throw_invalid("bvh.build: Cannot build a BVH with triangles but no vertices.");
        // Wide-node ids and per-level counts travel in 30-bit fields (look-back status words,
        // parent links with the slot in the low two bits).
        if (geometry->triangle_count > (1ull << 30) - 1ull)
            gwn::throw_invalid("bvh.build: Triangle count exceeds the native wide-collapse counter range.");
        gwn_bvh_handle staging;
        staging.stream = stream;
        gwn::build_bvh4(*geometry, staging, method, hploc_search_radius, stream);
        staging.storage_generation = gwn::next_storage_generation();
        if (geometry->triangle_count != 0) {
            staging.revision = gwn::reserve_revision();
            if (staging.revision == 0)
                gwn::throw_internal("bvh.build: BVH geometry-derived state revision is exhausted.");
        }
        std::swap(*bvh, staging);
        // staging now owns the replaced storage; its buffers free on their own streams.
    });
}

int gwn_refit_bvh_f32(gwn_geometry_t geometry, gwn_bvh_t bvh, hipStream_t stream) {
    return guarded("gwn_refit_bvh", [&] {
        if (geometry == nullptr || (geometry->triangle_count != 0 && geometry->i0.ptr == nullptr))
            gwn::throw_invalid("bvh.refit: Geometry accessor is invalid.");
        if (bvh == nullptr || !bvh->valid())
            gwn::throw_invalid("bvh.refit: BVH accessor is invalid.");
        if (geometry->triangle_count != bvh->triangle_count)
            gwn::throw_invalid("bvh.refit: Geometry triangle count differs from the BVH primitive order.");
        std::uint64_t const next_revision = gwn::reserve_revision();
        if (next_revision == 0)
            gwn::throw_internal("bvh.refit: BVH geometry-derived state revision is exhausted.");
        // Writes below are in place and cannot be rolled back: the BVH is unqueryable (and
        // every moment formed from it stale) until the refit has converged.  From here on its
        // storage is released on the refit stream, which holds the last writes.
        bvh->revision = 0;
        bvh->stream = stream;
        bvh->nodes.stream = bvh->primitive_indices.stream = bvh->triangles.stream = stream;
        bvh->parent.stream = bvh->internal_children.stream = stream;
        gwn::refit_bounds(*geometry, *bvh, stream);
        bvh->revision = next_revision;
    });
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
        info->width = 4;
        std::memcpy(info->root_bounds, &bvh->root.bounds, sizeof(info->root_bounds));
        info->root_reference = bvh->root.reference;
    });
}

int gwn_bvh_download(gwn_bvh_t bvh, void *nodes, uint32_t *primitive_indices, float *triangles, hipStream_t stream) {
    return guarded("gwn_bvh_download", [&] {
        require_bvh(bvh);
        if (nodes && bvh->node_count)
            GWN_HIP(hipMemcpyAsync(nodes, bvh->nodes.ptr, bvh->node_count * sizeof(gwn::bvh_node4), hipMemcpyDeviceToHost, stream));
        if (primitive_indices)
            GWN_HIP(hipMemcpyAsync(primitive_indices, bvh->primitive_indices.ptr, bvh->triangle_count * sizeof(std::uint32_t), hipMemcpyDeviceToHost, stream));
        if (triangles)
            GWN_HIP(hipMemcpyAsync(triangles, bvh->triangles.ptr, bvh->triangle_count * sizeof(gwn::tri_record), hipMemcpyDeviceToHost, stream));
        GWN_HIP(hipStreamSynchronize(stream));
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

int gwn_refit_bvh_moment_f32(gwn_bvh_t bvh, gwn_moment_t moment, hipStream_t stream) {
    return guarded("gwn_refit_bvh_moment", [&] {
        if (bvh == nullptr || !bvh->valid())
            gwn::throw_invalid("bvh.refit.moment: BVH accessor is invalid for moment refit.");
        if (moment == nullptr)
            gwn::throw_invalid("bvh.refit.moment: Moment handle is null.");
        gwn_moment_handle staging;
        staging.order = moment->order;
        staging.stream = stream;
        staging.storage_generation = gwn::next_storage_generation();
        gwn::refit_moment(*bvh, staging, stream);
        std::swap(*moment, staging);
    });
}

int gwn_moment_download(gwn_moment_t moment, float *fields, hipStream_t stream) {
    return guarded("gwn_moment_download", [&] {
        if (moment == nullptr || fields == nullptr)
            gwn::throw_invalid("Moment handle or output pointer is null.");
        std::size_t const N = moment->node_count;
        int const F = gwn::moment_field_count(moment->order);
        int const RV = gwn::moment_record_vec4(moment->order);
        if (N == 0 || moment->records.ptr == nullptr)
            return;
        std::vector<float> host(N * RV * 4);
        GWN_HIP(hipMemcpyAsync(host.data(), moment->records.ptr, host.size() * sizeof(float), hipMemcpyDeviceToHost, stream));
        GWN_HIP(hipStreamSynchronize(stream));
        for (std::size_t n = 0; n < N; ++n)
            std::memcpy(fields + n * F * 4, host.data() + n * RV * 4 + 4, sizeof(float) * F * 4);
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
        if (tuning->traversal != 0 && tuning->traversal != 1)
            gwn::throw_invalid("Batched query traversal must be 0 (per lane) or 1 (wave).");
        gwn::batch_set_tuning(*batch, tuning->block_size, tuning->chunk, tuning->blocks_per_cu,
                              tuning->traversal == 1, stream);
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
        if (gwn::batch_status(*batch, stream) != 0u)
            gwn::throw_internal("batch: Bottom-up moment refit did not converge.");
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
