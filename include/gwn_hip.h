/*
 * gwn_hip.h - C ABI of the MI355X (gfx950) geometric-query library libgwn_hip.so.
 *
 * Each entry point replaces one function of the reference's header-only gwn_* C++ API
 * (kririae/smallgwn, paths relative to its repository root).  Objects are opaque handles
 * that own stream-ordered device memory, like the reference's gwn_*_object classes; batch
 * queries take plain device pointers and a count, like its gwn_device_span arguments.
 *
 * Every function returns a gwn_error_code.  On failure gwn_last_error_message() returns the
 * diagnostic of the calling thread's last failure.  Status categories mirror
 * include/gwn/gwn_utils.cuh:213-218 (success, invalid_argument, cuda_runtime_error -> HIP
 * runtime error, internal_error).  Batch queries enqueue on the given stream and do not
 * synchronize (include/gwn/gwn_query.cuh:10-18).  Build and refit synchronize internally,
 * as the reference's do (include/gwn/detail/gwn_bvh_build_impl.cuh:230-240).
 *
 * The C++ template surface that mirrors the reference's names is include/gwn/gwn.hpp.
 */
#ifndef GWN_HIP_H
#define GWN_HIP_H

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GWN_HIP_ABI_VERSION 3

typedef enum gwn_error_code {
    GWN_SUCCESS = 0,
    GWN_ERROR_INVALID_ARGUMENT = 1,
    GWN_ERROR_HIP_RUNTIME = 2,
    GWN_ERROR_INTERNAL = 3
} gwn_error_code;

/* include/gwn/gwn_bvh_build.cuh:16-27 (enum order kept: H-PLOC first). */
typedef enum gwn_bvh_build_method {
    GWN_BVH_BUILD_HPLOC = 0,
    GWN_BVH_BUILD_LBVH = 1
} gwn_bvh_build_method;

/*
 * Arithmetic of a winding query.  GWN_PRECISION_PLAIN is the reference's float arithmetic.
 * GWN_PRECISION_SPLIT keeps traversal and the Taylor far field in FP32 and evaluates every
 * near-field (exact) triangle term in FP64.  GWN_PRECISION_FP64 evaluates every term in FP64
 * (the oracle mode of the exact batch).
 */
typedef enum gwn_precision {
    GWN_PRECISION_PLAIN = 0,
    GWN_PRECISION_SPLIT = 1,
    GWN_PRECISION_FP64 = 2
} gwn_precision;

typedef struct gwn_geometry_handle *gwn_geometry_t;
typedef struct gwn_bvh_handle *gwn_bvh_t;
typedef struct gwn_moment_handle *gwn_moment_t;
typedef struct gwn_boundary_handle *gwn_boundary_t;
typedef struct gwn_batch_handle *gwn_batch_t;

/* Reference gwn_query_batch_config (include/gwn/gwn_query.cuh:41-61): block size 256 and
 * stack capacity 64 by default.  sort_queries: -1 auto, 0 off, 1 on (queries are traversed in
 * Hilbert-curve order of their positions; results are written in input order).  The unsigned
 * distance batch also takes 2: Hilbert order with a wave-cooperative traversal (a wave walks each node once
 * for its 64 points, each point pruning by its own best distance).  Its distances are each
 * point's minimum over the triangles it reaches, which can differ from the reference order's in
 * the last FP32 bits where a box distance rounds above a contained triangle's distance.
 * lanes_per_point (unsigned distance batch only; other batches ignore it): 0 (default: four),
 * 1, or 4.  With four, the lanes of a quad share each BVH4 node, one child slot each; distances
 * and nearest records equal the one-lane traversal's bit for bit.  Ignored with
 * sort_queries = 2. */
typedef struct gwn_query_config {
    int32_t block_size;
    int32_t stack_capacity;
    int32_t sort_queries;
    int32_t lanes_per_point;
} gwn_query_config;

/* Host-side description of a built BVH (the scalar members of the reference's
 * gwn_bvh_accessor, include/gwn/gwn_bvh.cuh:209-223). */
typedef struct gwn_bvh_info {
    uint64_t node_count;
    uint64_t triangle_count;
    uint64_t revision;
    uint32_t internal_stack_bound;
    uint32_t packed_stack_bound;
    uint32_t max_depth;
    uint32_t width;
    float root_bounds[6];
    uint64_t root_reference;
} gwn_bvh_info;

/* Ray first-hit record, byte-compatible with gwn_ray_first_hit_result<float, uint32_t>
 * (include/gwn/gwn_query.cuh:63-95): t = -1 and primitive_id = 0xFFFFFFFF on a miss;
 * (u, v) weight the second and third vertex; the geometric normal is unnormalized. */
typedef enum gwn_ray_first_hit_status {
    GWN_RAY_MISS = 0,
    GWN_RAY_HIT = 1,
    GWN_RAY_OVERFLOW = 2
} gwn_ray_first_hit_status;

typedef struct gwn_ray_first_hit_result {
    float t, u, v;
    float geometric_normal_x, geometric_normal_y, geometric_normal_z;
    uint32_t primitive_id;
    uint8_t status;
    uint8_t padding[3];
} gwn_ray_first_hit_result;

int gwn_abi_version(void);
const char *gwn_last_error_message(void);

/* ---- accessors: non-owning snapshots of an object's device storage, replacing the reference's
 *      trivially copyable gwn_*_accessor structs (include/gwn/gwn_geometry.cuh:19-58,
 *      gwn_bvh.cuh:209-258, 260-329, gwn_boundary.cuh:28-54).  Any upload, build, refit, clear
 *      or destroy of the object invalidates them. ---- */
typedef struct gwn_geometry_accessor {
    const float *x, *y, *z;           /* vertex_count each */
    const uint32_t *i0, *i1, *i2;     /* triangle_count each */
    uint64_t vertex_count;
    uint64_t triangle_count;
} gwn_geometry_accessor;

typedef struct gwn_bvh_accessor {
    const void *nodes;                /* node_count gwn_bvh_node<4, float> (128 B each) */
    uint64_t node_count;
    const uint32_t *primitive_indices; /* leaf order -> original triangle id */
    const float *triangles;           /* leaf-ordered records, 9 floats each: v0, v1, v2 */
    uint64_t triangle_count;
    float root_bounds[6];
    uint64_t root_reference;          /* gwn_bvh_child reference bits of the root */
    uint32_t internal_stack_bound;
    uint32_t packed_stack_bound;
    uint64_t revision;                /* 0: not queryable */
} gwn_bvh_accessor;

typedef struct gwn_moment_accessor {
    const void *records;              /* node_count traversal records of record_bytes each */
    uint64_t node_count;
    int32_t order;
    uint32_t record_bytes;            /* 128, 256 or 384 for orders 0, 1, 2 */
    int32_t width;                    /* the width of the BVH the records were formed from */
    uint64_t bvh_revision;            /* the BVH revision the records were formed from */
    const void *eval_records;         /* node_count evaluation records (144, 272 or 400 B for
                                         orders 0, 1, 2): the coefficients the value-only Taylor
                                         queries read, one contiguous block per child slot */
} gwn_moment_accessor;

typedef struct gwn_boundary_accessor {
    const uint32_t *start, *end;      /* edge_count directed boundary rows */
    const uint64_t *multiplicity;
    uint64_t edge_count;
    uint64_t mesh_vertex_count;
    uint64_t mesh_triangle_count;
    int32_t is_built;
} gwn_boundary_accessor;

/* ---- geometry: replaces gwn_geometry_object + gwn_upload_geometry / gwn_update_geometry
 *      (include/gwn/gwn_geometry.cuh:19-149, detail/gwn_geometry_impl.cuh:102-219) ---- */
int gwn_geometry_create(gwn_geometry_t *out);
int gwn_geometry_destroy(gwn_geometry_t geometry);
/* Host SoA positions and oriented triangle indices; validates every index < vertex_count. */
int gwn_upload_geometry_f32(
    gwn_geometry_t geometry, const float *x, const float *y, const float *z, uint64_t vertex_count,
    const uint32_t *i0, const uint32_t *i1, const uint32_t *i2, uint64_t triangle_count,
    hipStream_t stream);
/* Replace positions (same vertex count) from host memory.  Enqueued on `stream` without a host
 * synchronization (include/gwn/detail/gwn_geometry_impl.cuh:172-178): pageable sources are staged
 * before the call returns, pinned sources must stay valid until the stream reaches the copies. */
int gwn_update_geometry_f32(
    gwn_geometry_t geometry, const float *x, const float *y, const float *z, uint64_t vertex_count,
    hipStream_t stream);
/* Replace positions from device memory (enqueued device-to-device copies, no host sync). */
int gwn_update_geometry_device_f32(
    gwn_geometry_t geometry, const float *x, const float *y, const float *z, uint64_t vertex_count,
    hipStream_t stream);
int gwn_geometry_counts(gwn_geometry_t geometry, uint64_t *vertex_count, uint64_t *triangle_count);

/* ---- Real = double and Index = uint64: gwn_geometry_object<Real, Index> for the reference's other
 *      template arguments (include/gwn/gwn_geometry.cuh:19-58, gwn_bvh_build.cuh:34-39).
 *      uint64 indices are validated like uint32 ones and stored as uint32, so vertex and triangle
 *      counts stay below 2^32.  A Real = double geometry keeps its positions in double (a float
 *      rounding of them steers the build's Morton codes only) and is built, refit, moment-refit
 *      and queried through the _f64 entry points below; the _f32 entry points reject it, and the
 *      _f64 ones reject Real = float objects.  Device accessors are provided for Real = float. ---- */
int gwn_upload_geometry_f32_u64(
    gwn_geometry_t geometry, const float *x, const float *y, const float *z, uint64_t vertex_count,
    const uint64_t *i0, const uint64_t *i1, const uint64_t *i2, uint64_t triangle_count,
    hipStream_t stream);
int gwn_upload_geometry_f64(
    gwn_geometry_t geometry, const double *x, const double *y, const double *z, uint64_t vertex_count,
    const uint32_t *i0, const uint32_t *i1, const uint32_t *i2, uint64_t triangle_count,
    hipStream_t stream);
int gwn_upload_geometry_f64_u64(
    gwn_geometry_t geometry, const double *x, const double *y, const double *z, uint64_t vertex_count,
    const uint64_t *i0, const uint64_t *i1, const uint64_t *i2, uint64_t triangle_count,
    hipStream_t stream);
/* As gwn_update_geometry_f32 / gwn_update_geometry_device_f32, for a Real = double geometry. */
int gwn_update_geometry_f64(
    gwn_geometry_t geometry, const double *x, const double *y, const double *z, uint64_t vertex_count,
    hipStream_t stream);
int gwn_update_geometry_device_f64(
    gwn_geometry_t geometry, const double *x, const double *y, const double *z, uint64_t vertex_count,
    hipStream_t stream);

/* ---- object stream binding and release (include/gwn/gwn_utils.cuh:59-80 gwn_stream_mixin;
 *      gwn_geometry.cuh:117-125, gwn_bvh.cuh:356-415, gwn_boundary.cuh:84-110 clear()).  An object
 *      is bound to the stream of the last operation that replaced or wrote its storage; its
 *      storage is released on that bound stream (by clear, destroy, or a replacement made on
 *      another stream), so work already queued there on the old storage finishes first.
 *      set_stream rebinds without synchronizing either stream; clear releases on the current
 *      binding and keeps it.  get_accessor returns a non-owning snapshot of the device storage.
 *      The update and refit calls above rebind to their stream, as the reference's do. ---- */
int gwn_geometry_stream(gwn_geometry_t geometry, hipStream_t *stream);
int gwn_geometry_set_stream(gwn_geometry_t geometry, hipStream_t stream);
int gwn_geometry_clear(gwn_geometry_t geometry);
int gwn_geometry_get_accessor(gwn_geometry_t geometry, gwn_geometry_accessor *accessor);
int gwn_bvh_stream(gwn_bvh_t bvh, hipStream_t *stream);
int gwn_bvh_set_stream(gwn_bvh_t bvh, hipStream_t stream);
int gwn_bvh_clear(gwn_bvh_t bvh);
int gwn_bvh_get_accessor(gwn_bvh_t bvh, gwn_bvh_accessor *accessor);
int gwn_moment_stream(gwn_moment_t moment, hipStream_t *stream);
int gwn_moment_set_stream(gwn_moment_t moment, hipStream_t stream);
int gwn_moment_clear(gwn_moment_t moment);
int gwn_moment_get_accessor(gwn_moment_t moment, gwn_moment_accessor *accessor);
int gwn_boundary_stream(gwn_boundary_t boundary, hipStream_t *stream);
int gwn_boundary_set_stream(gwn_boundary_t boundary, hipStream_t stream);
int gwn_boundary_clear(gwn_boundary_t boundary);
int gwn_boundary_get_accessor(gwn_boundary_t boundary, gwn_boundary_accessor *accessor);

/* ---- BVH: replaces gwn_bvh_object<4, float, uint32_t> and gwn_build_bvh
 *      (include/gwn/gwn_bvh.cuh:331-380, include/gwn/gwn_bvh_build.cuh:29-39) ---- */
int gwn_bvh_create(gwn_bvh_t *out);
int gwn_bvh_destroy(gwn_bvh_t bvh);
/* Width-4 build.  hploc_search_radius must be in [1, 8] for H-PLOC (build_impl.cuh:66-72). */
int gwn_build_bvh4_f32(
    gwn_geometry_t geometry, gwn_bvh_t bvh, int32_t method, uint32_t hploc_search_radius,
    hipStream_t stream);
/* Build of width 2, 3 or 4 (gwn_bvh_object<Width, float, uint32_t>, include/gwn/gwn_bvh.cuh:331-380;
 * the collapse expands each node up to Width children, detail/gwn_bvh_build_collapse.cuh:163-296).
 * Narrower trees are stored four-wide with their extra slots empty, so every refit and query
 * below accepts them; gwn_bvh_download and gwn_moment_download return the reference's
 * Width-wide layouts.  gwn_build_bvh4_f32 is width 4. */
int gwn_build_bvh_f32(
    gwn_geometry_t geometry, gwn_bvh_t bvh, int32_t width, int32_t method, uint32_t hploc_search_radius,
    hipStream_t stream);
/* Bounds refit from updated positions (same triangle indexing): replaces gwn_refit_bvh
 * (include/gwn/gwn_bvh_refit.cuh:25-29, detail/gwn_bvh_refit_impl.cuh:194-320).  Rewrites the
 * leaf-ordered triangle records and every child box in place, keeps topology, primitive order
 * and stack bounds, and issues a new revision, so moments must be refit afterwards.  A failure
 * after the in-place writes began leaves the BVH unqueryable (revision 0); it is then bound
 * to `stream` like a successful refit. */
int gwn_refit_bvh_f32(gwn_geometry_t geometry, gwn_bvh_t bvh, hipStream_t stream);
/* gwn_bvh_object<Width, double, Index> build and refit from a Real = double geometry: the tree is
 * built from the float-rounded positions, then every leaf record and child box is recomputed in
 * double from the double positions, one launch per breadth-first level, deepest first
 * (detail/gwn_bvh_refit_impl.cuh:23-352 with Real = double).  Boxes are exact unions of the
 * double triangles; the tree shape alone depends on the rounding. */
int gwn_build_bvh_f64(
    gwn_geometry_t geometry, gwn_bvh_t bvh, int32_t width, int32_t method, uint32_t hploc_search_radius,
    hipStream_t stream);
int gwn_refit_bvh_f64(gwn_geometry_t geometry, gwn_bvh_t bvh, hipStream_t stream);
int gwn_bvh_get_info(gwn_bvh_t bvh, gwn_bvh_info *info);
/* Copies the canonical arrays to host memory (any pointer may be NULL) and synchronizes:
 * nodes in the reference's gwn_bvh_node<Width, float> layout (32 B per child), primitive order,
 * and leaf-ordered triangle records as 9 floats (v0, v1, v2). */
int gwn_bvh_download(
    gwn_bvh_t bvh, void *nodes, uint32_t *primitive_indices, float *triangles, hipStream_t stream);
/* Real = double: gwn_bvh_node<Width, double> (56 B per child: six double bounds, then the
 * 64-bit reference) and records as 9 doubles. */
int gwn_bvh_download_f64(
    gwn_bvh_t bvh, void *nodes, uint32_t *primitive_indices, double *triangles, hipStream_t stream);

/* ---- boundary chain: replaces gwn_boundary_chain_object<uint32_t> and gwn_build_boundary_chain
 *      (include/gwn/gwn_boundary.cuh:28-125, detail/gwn_boundary_impl.cuh:21-413).  Directed rows
 *      (start, end, multiplicity) of the algebraic boundary, in (min vertex, max vertex) order;
 *      a closed mesh builds an empty chain.  Strong guarantee (staging + swap). ---- */
int gwn_boundary_create(gwn_boundary_t *out);
int gwn_boundary_destroy(gwn_boundary_t boundary);
int gwn_build_boundary_chain(gwn_geometry_t geometry, gwn_boundary_t boundary, hipStream_t stream);
int gwn_boundary_info(
    gwn_boundary_t boundary, uint64_t *edge_count, uint64_t *mesh_vertex_count, uint64_t *mesh_triangle_count,
    int32_t *is_built);
/* Copies the rows to host memory (any pointer may be NULL) and synchronizes. */
int gwn_boundary_download(
    gwn_boundary_t boundary, uint32_t *start, uint32_t *end, uint64_t *multiplicity, hipStream_t stream);

/* ---- Taylor moments: replaces gwn_bvh_moment_object<4, Order, float> and
 *      gwn_refit_bvh_moment<Order> (include/gwn/gwn_bvh_refit.cuh:31-51) ---- */
int gwn_moment_create(gwn_moment_t *out, int32_t order);
int gwn_moment_destroy(gwn_moment_t moment);
int gwn_refit_bvh_moment_f32(gwn_bvh_t bvh, gwn_moment_t moment, hipStream_t stream);
/* Copies coefficients in the reference's field-SoA gwn_bvh_moment_node<Width, Order, float>
 * layout (node-major, then field, then child slot; Width of the BVH they were formed from) and
 * synchronizes. */
int gwn_moment_download(gwn_moment_t moment, float *fields, hipStream_t stream);
int gwn_moment_order(gwn_moment_t moment, int32_t *order);
/* Node count and the BVH revision the moments were formed from, for either Real (the object
 * state behind gwn_bvh_moment_object::is_valid_for, include/gwn/gwn_bvh.cuh:286-293). */
int gwn_moment_info(gwn_moment_t moment, uint64_t *node_count, uint64_t *bvh_revision);
/* gwn_bvh_moment_object<Width, Order, double> from a Real = double BVH: the same per-triangle
 * moments, sums in child-slot order and recentring as the float refit, in double, one launch
 * per level (detail/gwn_bvh_refit_moment_impl.cuh:17-279). */
int gwn_refit_bvh_moment_f64(gwn_bvh_t bvh, gwn_moment_t moment, hipStream_t stream);
int gwn_moment_download_f64(gwn_moment_t moment, double *fields, hipStream_t stream);

/* ---- winding-number batches: replace gwn_compute_winding_number_exact_batch and
 *      gwn_compute_winding_number_taylor_batch<Order> (include/gwn/gwn_query.cuh:308-445) ---- */
int gwn_compute_winding_number_exact_batch_f32(
    gwn_bvh_t bvh, const float *qx, const float *qy, const float *qz, float *out, uint64_t count,
    const gwn_query_config *config, hipStream_t stream);
int gwn_compute_winding_number_taylor_batch_f32(
    gwn_bvh_t bvh, gwn_moment_t moment, const float *qx, const float *qy, const float *qz,
    float *out, uint64_t count, float accuracy_scale, const gwn_query_config *config,
    hipStream_t stream);
/* Real = double (the same batches instantiated with double, gwn_query.cuh:308-445): the exact sum
 * of the double Van Oosterom-Strackee solid angles, and the Taylor traversal with the acceptance
 * test and far-field terms in double.  One query per lane, in input order. */
int gwn_compute_winding_number_exact_batch_f64(
    gwn_bvh_t bvh, const double *qx, const double *qy, const double *qz, double *out, uint64_t count,
    const gwn_query_config *config, hipStream_t stream);
int gwn_compute_winding_number_taylor_batch_f64(
    gwn_bvh_t bvh, gwn_moment_t moment, const double *qx, const double *qy, const double *qz,
    double *out, uint64_t count, double accuracy_scale, const gwn_query_config *config,
    hipStream_t stream);

/* Precision-split and FP64 forms (no reference counterpart: the reference computes in the
 * geometry's Real).  The exact form writes double results for every precision. */
int gwn_compute_winding_number_exact_batch_ex_f32(
    gwn_bvh_t bvh, const float *qx, const float *qy, const float *qz, double *out, uint64_t count,
    int32_t precision, const gwn_query_config *config, hipStream_t stream);
/* order (optional, device, count entries): visit queries in this permutation (from
 * gwn_compute_query_order_f32); when NULL the batch presorts by itself as config says.
 * counters (optional, device, 3 x uint32 per query): child tests, accepted expansions,
 * near-field triangle evaluations -- the work counts used for the FLOP roofline. */
int gwn_compute_winding_number_taylor_batch_ex_f32(
    gwn_bvh_t bvh, gwn_moment_t moment, const float *qx, const float *qy, const float *qz,
    float *out, uint64_t count, float accuracy_scale, int32_t precision,
    const gwn_query_config *config, const uint32_t *order, uint32_t *counters, hipStream_t stream);

/* Allocation-free Taylor winding batch (no reference counterpart).  For callers that cannot
 * allocate device memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  Same arguments, validation and results as
 * gwn_compute_winding_number_taylor_batch_ex_f32 without an order or counters.  The Hilbert
 * presort's scratch lives in a static per-GPU device array (about 16.3 MiB, held until the
 * process ends) instead of stream-ordered allocations, so count is limited to 1048576 queries
 * (larger batches return GWN_ERROR_INVALID_ARGUMENT; split them).  The array is shared by every
 * call on one GPU: calls that overlap on one GPU (other streams or host threads) race. */
int gwn_compute_winding_number_taylor_batch_noalloc_f32(
    gwn_bvh_t bvh, gwn_moment_t moment, const float *qx, const float *qy, const float *qz,
    float *out, uint64_t count, float accuracy_scale, int32_t precision,
    const gwn_query_config *config, hipStream_t stream);

/* ---- gradient batch: replaces gwn_compute_winding_gradient_taylor_batch<Order, Config>
 *      (include/gwn/gwn_query.cuh:684-756, detail/gwn_query_winding_impl.cuh:20-99, 217-325).
 *      Analytic gradient of the Taylor winding number, written to gx, gy, gz. ---- */
int gwn_compute_winding_gradient_taylor_batch_f32(
    gwn_bvh_t bvh, gwn_moment_t moment, const float *qx, const float *qy, const float *qz, float *gx,
    float *gy, float *gz, uint64_t count, float accuracy_scale, const gwn_query_config *config,
    hipStream_t stream);
/* Winding number (out, optional) and gradient from one traversal, with the precision split, a
 * caller-given query order and the work counters of gwn_compute_winding_number_taylor_batch_ex_f32. */
int gwn_compute_winding_gradient_taylor_batch_ex_f32(
    gwn_bvh_t bvh, gwn_moment_t moment, const float *qx, const float *qy, const float *qz, float *out,
    float *gx, float *gy, float *gz, uint64_t count, float accuracy_scale, int32_t precision,
    const gwn_query_config *config, const uint32_t *order, uint32_t *counters, hipStream_t stream);

/* ---- unsigned distance: replaces gwn_compute_unsigned_distance_batch<Config>
 *      (include/gwn/gwn_query.cuh:129-180, detail/gwn_query_distance_impl.cuh:36-192).
 *      culling_band is the initial bound and the result when no triangle is closer (negative
 *      values clamp to 0, +inf searches unbounded); a query whose traversal would exceed the
 *      stack capacity gets NaN. ---- */
int gwn_compute_unsigned_distance_batch_f32(
    gwn_bvh_t bvh, const float *qx, const float *qy, const float *qz, float *out, uint64_t count,
    float culling_band, const gwn_query_config *config, hipStream_t stream);
/* The same, also writing each query's nearest triangle (no reference counterpart; the input of
 * gwn_compute_unsigned_distance_adjoint_batch_ex_f32): its BVH leaf record, or 0xFFFFFFFF when the
 * distance is the culling band or zero (no gradient), 0xFFFFFFFE on stack overflow. */
int gwn_compute_unsigned_distance_batch_ex_f32(
    gwn_bvh_t bvh, const float *qx, const float *qy, const float *qz, float *out, uint32_t *nearest,
    uint64_t count, float culling_band, const gwn_query_config *config, hipStream_t stream);

/* ---- ray first hit: replaces gwn_compute_ray_first_hit_batch<Config>
 *      (include/gwn/gwn_query.cuh:205-290, detail/gwn_query_ray_impl.cuh:20-595).  Rays are
 *      tested on the closed interval [t_min, t_max]; the traversal is the reference's packed
 *      near-first one when packed_stack_bound fits the stack capacity, else its canonical one,
 *      so ties between equal-t hits resolve in the reference's order. ---- */
int gwn_compute_ray_first_hit_batch_f32(
    gwn_bvh_t bvh, const float *ox, const float *oy, const float *oz, const float *dx, const float *dy,
    const float *dz, gwn_ray_first_hit_result *out, uint64_t count, float t_min, float t_max,
    const gwn_query_config *config, hipStream_t stream);

/* ---- traversal work counters (no reference counterpart; roofline accounting).  Two uint32 per
 *      query: child bounding-box tests and triangle tests of the one-lane unsigned-distance
 *      traversal (which the four-lanes-per-point form replays step for step), or slab tests and
 *      ray-triangle tests of the ray batch's traversal.  No distances or hits are written. ---- */
int gwn_compute_unsigned_distance_work_f32(
    gwn_bvh_t bvh, const float *qx, const float *qy, const float *qz, uint64_t count, float culling_band,
    uint32_t *counters, hipStream_t stream);
int gwn_compute_ray_first_hit_work_f32(
    gwn_bvh_t bvh, const float *ox, const float *oy, const float *oz, const float *dx, const float *dy,
    const float *dz, uint64_t count, float t_min, float t_max, const gwn_query_config *config, uint32_t *counters,
    hipStream_t stream);

/* ---- antipodal winding number: replaces gwn_compute_winding_number_antipodal_batch<Config>
 *      (include/gwn/gwn_query.cuh:492-552, detail/gwn_query_winding_antipodal_impl.cuh:55-490).
 *      Signed crossings of the positive coordinate ray with the BVH's triangles plus the boundary
 *      chain's spherical areas over 4 pi; a singular axis retries in the order +Z, +X, +Y and a
 *      query singular on all three (or whose traversal exceeds the stack) gets NaN.  The boundary
 *      chain must come from geometry's topology and the BVH from the same geometry. ---- */
int gwn_compute_winding_number_antipodal_batch_f32(
    gwn_geometry_t geometry, gwn_bvh_t bvh, gwn_boundary_t boundary, const float *qx, const float *qy,
    const float *qz, float *out, uint64_t count, const gwn_query_config *config, hipStream_t stream);
/* Precision split (no reference counterpart): classification and crossings in FP32, each
 * boundary area in FP64 on the exact vertices, summed in FP64. */
int gwn_compute_winding_number_antipodal_batch_ex_f32(
    gwn_geometry_t geometry, gwn_bvh_t bvh, gwn_boundary_t boundary, const float *qx, const float *qy,
    const float *qz, float *out, uint64_t count, int32_t precision, const gwn_query_config *config,
    hipStream_t stream);

/* ---- antipodal gradient: replaces gwn_compute_winding_gradient_antipodal_batch<Config>
 *      (include/gwn/gwn_query.cuh:582-640, detail/gwn_query_winding_antipodal_impl.cuh:257-376,
 *      492-552).  Gradient of the boundary area term; (NaN, NaN, NaN) when all axes are
 *      singular.  config->stack_capacity is not used. ---- */
int gwn_compute_winding_gradient_antipodal_batch_f32(
    gwn_geometry_t geometry, gwn_boundary_t boundary, const float *qx, const float *qy, const float *qz,
    float *gx, float *gy, float *gz, uint64_t count, const gwn_query_config *config, hipStream_t stream);
int gwn_compute_winding_gradient_antipodal_batch_ex_f32(
    gwn_geometry_t geometry, gwn_boundary_t boundary, const float *qx, const float *qy, const float *qz,
    float *gx, float *gy, float *gz, uint64_t count, int32_t precision, const gwn_query_config *config,
    hipStream_t stream);

/* ---- adjoint batches (no reference counterpart; the reference differentiates only with respect
 *      to the query point, include/gwn/gwn_query.cuh:656-756).  For a loss L with upstream
 *      gradients grad_w[q] = dL/dw_q (Taylor winding, orders 0 and 1, with the forward batch's
 *      accept/descend decisions for the same accuracy_scale) or grad_d[p] = dL/dd_p (unsigned
 *      distance), the batches ADD dL/dV to grad_vx/vy/vz (vertex_count floats each, in the
 *      geometry's vertex order) and write dL/dq to grad_qx/qy/qz (or dL/dp) when those are non-null.
 *      The BVH and moments must describe the geometry's current positions (refit after an update).
 *      Sums use FP32 atomics: results are reproducible to rounding, not bit for bit. ---- */
int gwn_compute_winding_taylor_adjoint_batch_f32(
    gwn_geometry_t geometry, gwn_bvh_t bvh, gwn_moment_t moment, const float *qx, const float *qy,
    const float *qz, const float *grad_w, uint64_t count, float accuracy_scale, float *grad_vx, float *grad_vy,
    float *grad_vz, float *grad_qx, float *grad_qy, float *grad_qz, const gwn_query_config *config,
    hipStream_t stream);
int gwn_compute_unsigned_distance_adjoint_batch_f32(
    gwn_geometry_t geometry, gwn_bvh_t bvh, const float *px, const float *py, const float *pz,
    const float *grad_d, uint64_t count, float culling_band, float *grad_vx, float *grad_vy, float *grad_vz,
    float *grad_px, float *grad_py, float *grad_pz, const gwn_query_config *config, hipStream_t stream);
/* The same from the forward's nearest records (gwn_compute_unsigned_distance_batch_ex_f32 on the
 * same BVH and points): no traversal runs. */
int gwn_compute_unsigned_distance_adjoint_batch_ex_f32(
    gwn_geometry_t geometry, gwn_bvh_t bvh, const float *px, const float *py, const float *pz,
    const uint32_t *nearest, const float *grad_d, uint64_t count, float *grad_vx, float *grad_vy, float *grad_vz,
    float *grad_px, float *grad_py, float *grad_pz, hipStream_t stream);

/* ---- batched mode (no reference counterpart; the reference runs one mesh per call).  A batch
 *      fixes, for each mesh, its BVH and moment objects (all of one Taylor order) and its query
 *      and output device pointers.  gwn_batch_create allocates everything a replay needs and
 *      refits the moments once (synchronizes).  gwn_batch_refit_moments enqueues every mesh's
 *      moment refit and gwn_batch_query one persistent Taylor launch over all meshes' queries
 *      (plain precision); neither allocates nor synchronizes, so both may be captured:
 *      gwn_batch_capture records [refits +] counter reset + query launch into a hipGraph on
 *      the batch's own stream, gwn_batch_replay launches it.  Results equal the single-mesh
 *      Taylor batch's bit for bit.  gwn_batch_tune times candidate launch shapes and keeps the
 *      fastest (synchronizes; drops a captured graph).  The objects and pointers must outlive
 *      the batch; replays of one batch must not overlap.  Each mesh needs its own moment object.
 *      A batch refuses to run once a mesh's BVH or moment storage was replaced (rebuild, upload,
 *      gwn_refit_bvh_moment_f32), and a query without refits (gwn_batch_query, or replaying a graph
 *      captured without them) refuses moments not formed from their BVH's current revision.
 *      gwn_batch_set_tuning and gwn_batch_tune first wait for the batch's last enqueued work. ---- */
typedef struct gwn_batch_item {
    gwn_bvh_t bvh;
    gwn_moment_t moment;
    const float *qx, *qy, *qz;
    float *out;
    uint64_t count;
} gwn_batch_item;

typedef struct gwn_batch_tuning {
    int32_t block_size;      /* 64, 128, 192 or 256 threads */
    int32_t chunk;           /* queries per work item, a multiple of block_size */
    int32_t blocks_per_cu;   /* persistent grid = blocks_per_cu x compute units */
    int32_t traversal;       /* 0: one query per lane; 1: wave-cooperative over each mesh's
                                Hilbert query order (computed at create); 2: per mesh, wave-
                                cooperative where the mesh has at least wave_queries_per_triangle
                                queries per triangle, one query per lane elsewhere.  All three
                                give the same results. */
    int32_t stack_capacity;  /* compiled traversal stack: 32, 64 or 128, at least every tree's
                                internal_stack_bound; 0 in gwn_batch_set_tuning keeps it */
    float wave_queries_per_triangle; /* traversal 2's density threshold (positive, finite); read
                                only for traversal 2.  It fills what was padding, so the layout
                                of every other field is unchanged. */
    uint64_t meshes;         /* read only */
    uint64_t queries;        /* read only */
} gwn_batch_tuning;

int gwn_batch_create(const gwn_batch_item *items, uint64_t item_count, float accuracy_scale, gwn_batch_t *out,
                     hipStream_t stream);
int gwn_batch_destroy(gwn_batch_t batch);
int gwn_batch_refit_moments(gwn_batch_t batch, hipStream_t stream);
int gwn_batch_query(gwn_batch_t batch, hipStream_t stream);
int gwn_batch_set_tuning(gwn_batch_t batch, const gwn_batch_tuning *tuning, hipStream_t stream);
int gwn_batch_get_tuning(gwn_batch_t batch, gwn_batch_tuning *tuning);
int gwn_batch_tune(gwn_batch_t batch, hipStream_t stream, float *best_ms);
int gwn_batch_capture(gwn_batch_t batch, int32_t include_refit);
int gwn_batch_replay(gwn_batch_t batch, hipStream_t stream);
/* Synchronizes `stream` and reports whether any moment refit or segmented build of the batch
 * (eager or in a graph replay) failed since the batch was made or since the last
 * gwn_batch_status call: GWN_ERROR_INTERNAL with "batch: Bottom-up moment refit did not
 * converge.", "batch: A segmented BVH build failed.", or (a batch over geometries whose updated
 * geometry gave a tree past what its launches were planned for) a message naming the level
 * launches or the stack capacity; the word is then cleared.  Else GWN_SUCCESS.  Refits and builds
 * enqueue without reading the word back, so they can be captured. */
int gwn_batch_status(gwn_batch_t batch, hipStream_t stream);

/* ---- batched mode over geometries (no reference counterpart).  The batch builds and owns every
 *      mesh's BVH (LBVH, or H-PLOC with the given search radius; the given width) and Taylor
 *      moments of the given order: all meshes' builds and moment refits run in one segmented pass
 *      of a fixed sequence of launches (scene bounds, Morton codes, one stable sort by (mesh,
 *      code), H-PLOC over all sorted leaves or Karras LBVH + one bounds launch per binary level,
 *      one plan + one persistent launch per wide level of the collapse, the level refit of the
 *      moments) that allocates nothing and never reads back, so it can be captured.
 *      Each mesh's tree, records and moments equal gwn_build_bvh_f32 + gwn_refit_bvh_moment_f32
 *      of that mesh alone, and its query results equal the single-mesh Taylor batch's, bit for bit.
 *      gwn_batch_create_built builds once (synchronizes) and plans the pass: the number of level
 *      launches (the deepest tree + 8) and the traversal stack (the largest bound + 8).  gwn_batch_build enqueues the
 *      pass; gwn_batch_capture(batch, 1) records pass + query launch, so a replay rebuilds every tree
 *      from the geometries' current positions (gwn_update_geometry_f32 between replays).  A tree
 *      past the plan sets the status word (gwn_batch_status); gwn_batch_set_build re-plans.
 *      gwn_batch_tune picks the query launch shape and stack; gwn_batch_tune_build times the pass
 *      plus the queries over widths {4, 3, 2} x {LBVH, H-PLOC radii 8, 4, 2, 1} and keeps the fastest
 *      (synchronizes).  gwn_batch_export copies one mesh's tree and moments into BVH and moment
 *      objects for the single-mesh entry points.  gwn_batch_refit_moments does not apply to such
 *      a batch.  The geometry objects (Real = float) must outlive the batch and keep their storage
 *      (gwn_update_geometry_f32 keeps it; an upload or clear replaces it and the batch refuses to
 *      run).  Device memory held: about 0.9 KB per triangle at order 1 (node slots for the
 *      widest case, moments and build scratch). ---- */
typedef struct gwn_batch_build_item {
    gwn_geometry_t geometry;
    const float *qx, *qy, *qz;
    float *out;
    uint64_t count;
} gwn_batch_build_item;

typedef struct gwn_batch_build_config {
    int32_t method;                /* GWN_BVH_BUILD_HPLOC or GWN_BVH_BUILD_LBVH */
    int32_t width;                 /* 2, 3 or 4 */
    uint32_t hploc_search_radius;  /* 1..8 (H-PLOC) */
    int32_t order;                 /* Taylor order 0, 1 or 2 (fixed at create) */
} gwn_batch_build_config;

int gwn_batch_create_built(const gwn_batch_build_item *items, uint64_t item_count, const gwn_batch_build_config *config,
                           float accuracy_scale, gwn_batch_t *out, hipStream_t stream);
int gwn_batch_build(gwn_batch_t batch, hipStream_t stream);
int gwn_batch_set_build(gwn_batch_t batch, const gwn_batch_build_config *config, hipStream_t stream);
int gwn_batch_get_build(gwn_batch_t batch, gwn_batch_build_config *config, int32_t *level_launches);
int gwn_batch_tune_build(gwn_batch_t batch, hipStream_t stream, float *best_ms);
int gwn_batch_export(gwn_batch_t batch, uint64_t mesh, gwn_bvh_t bvh, gwn_moment_t moment, hipStream_t stream);

/* Hilbert-curve-order permutation of query points within the BVH root bounds (positions on the
 * 1024^3 grid over the root box, 30-bit keys, stable radix sort): order[i] is the input index of
 * the i-th query to visit.  Reusable across
 * batches over the same points (no reference counterpart; an MI355X traversal aid). */
int gwn_compute_query_order_f32(
    gwn_bvh_t bvh, const float *qx, const float *qy, const float *qz, uint32_t *order,
    uint64_t count, hipStream_t stream);

/* ---- internal building block exported for tests: stable LSD radix sort of device
 *      (uint64 key, uint32 value) pairs on bits [begin_bit, end_bit) (replaces the CUB
 *      DeviceRadixSort::SortPairs call of include/gwn/detail/gwn_bvh_build_common.cuh:325-347).
 *      Sorted output overwrites keys/values; scratch is allocated stream-ordered. ---- */
int gwn_radix_sort_pairs_u64(
    uint64_t *keys, uint32_t *values, uint64_t count, int32_t begin_bit, int32_t end_bit,
    hipStream_t stream);
int gwn_radix_sort_pairs_u32(
    uint32_t *keys, uint32_t *values, uint64_t count, int32_t begin_bit, int32_t end_bit,
    hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* GWN_HIP_H */
