/*
 * gwn_oracle.h - CPU oracle for the smallgwn geometric-query path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation in smallgwn_amd/ is compared against.  Only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it;
 * the product path never calls into it.
 *
 * It is a plain-C restatement of the reference algorithms (citations are
 * relative to the reference repository root):
 *   - Morton encoding: include/gwn/detail/gwn_bvh_build_common.cuh:35-88, 157-214
 *   - scene bounds / safe inverse: gwn_bvh_build_common.cuh:243-274
 *   - stable key sort: gwn_bvh_build_common.cuh:325-347 (CUB LSD radix sort is stable)
 *   - Karras LBVH: include/gwn/detail/gwn_bvh_build_binary.cuh:29-129
 *   - bottom-up bounds: gwn_bvh_build_binary.cuh:131-192
 *   - bounds refit: include/gwn/detail/gwn_bvh_refit_impl.cuh:23-192
 *   - BVH4 collapse rule and stack bounds: include/gwn/detail/gwn_bvh_build_collapse.cuh:163-296
 *   - leaf-ordered triangle records: include/gwn/detail/gwn_bvh_triangle_impl.cuh:23-67
 *   - Taylor moments: include/gwn/detail/gwn_bvh_refit_common.cuh:66-387, 445-607
 *   - moment accumulation order: include/gwn/detail/gwn_bvh_refit_moment_impl.cuh:33-47
 *   - solid angle: include/gwn/detail/gwn_query_geometry_impl.cuh:16-45
 *   - exact winding: include/gwn/detail/gwn_query_winding_impl.cuh:377-389
 *   - Taylor traversal: include/gwn/detail/gwn_query_winding_impl.cuh:102-374
 *   - solid-angle gradient and Taylor gradient terms: gwn_query_winding_impl.cuh:20-99, 217-325
 *   - ray first hit: include/gwn/detail/gwn_query_ray_impl.cuh:20-595, gwn_query.cuh:214-290
 *   - boundary chain: include/gwn/detail/gwn_boundary_impl.cuh:21-354
 *   - antipodal winding and gradient: include/gwn/detail/gwn_query_winding_antipodal_impl.cuh:55-552
 *   - unsigned distance: include/gwn/detail/gwn_query_distance_impl.cuh:20-192,
 *     gwn_query_geometry_impl.cuh:46-114
 *
 * Deliberate, documented differences from the reference that the HIP path
 * shares:
 *   - wide nodes are numbered breadth-first by (depth, first sorted leaf of
 *     the subtree) instead of (depth, nondeterministic parent slot); the node
 *     set, child slot order, bounds and stack bounds are unchanged.
 *   - triangle records store the three vertices (v0, v1, v2).  The plain
 *     ("reference form") arithmetic reconstructs e1 = v1 - v0,
 *     b = v0 + e1 exactly as the reference's (v0, e1, e2) record does.
 *
 * Everything is compiled with -ffp-contract=off so float results are
 * reproducible operation by operation.
 */
#ifndef GWN_ORACLE_H
#define GWN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reference gwn_bvh_child<float> layout (include/gwn/gwn_bvh.cuh:55-103): 32 bytes. */
typedef struct {
    float bounds[6];      /* min_x min_y min_z max_x max_y max_z */
    uint64_t reference;   /* offset bits 0-46, valid bit 47, primitive count bits 48-63 */
} gwn_oracle_child_f32;

typedef struct {
    gwn_oracle_child_f32 children[4];
} gwn_oracle_node4_f32;

/* Binary LBVH node: kind 1 = internal, 2 = leaf (gwn_bvh_build_binary.cuh:13-27). */
typedef struct {
    uint32_t left_index;
    uint32_t right_index;
    uint32_t left_kind;
    uint32_t right_kind;
} gwn_oracle_binary_node;

typedef struct {
    /* caller-allocated outputs; capacities: nodes >= max(T-1,1), the rest >= T */
    gwn_oracle_node4_f32 *nodes;
    uint32_t *primitive_indices;
    float *triangles;              /* 9 floats per record: v0 xyz, v1 xyz, v2 xyz */
    uint64_t *sorted_morton;
    gwn_oracle_binary_node *binary_nodes; /* optional (may be NULL): T-1 entries */
    float *binary_bounds;          /* optional: 6 floats per binary internal node */
    /* scalar outputs */
    gwn_oracle_child_f32 root;
    uint64_t node_count;
    uint32_t internal_stack_bound;
    uint32_t packed_stack_bound;
    float scene[9];                /* min xyz, max xyz, inv xyz */
    uint32_t *binary_first;        /* optional: smallest sorted leaf under each binary internal node */
    uint32_t binary_root;          /* binary internal root (0 for LBVH) */
} gwn_oracle_bvh_f32;

/* Binary builder selection, gwn_bvh_build_options (gwn_bvh_build.cuh:16-27). */
enum { GWN_ORACLE_BUILD_HPLOC = 0, GWN_ORACLE_BUILD_LBVH = 1 };

enum {
    GWN_ORACLE_OK = 0,
    GWN_ORACLE_INVALID_ARGUMENT = 1,
    GWN_ORACLE_INTERNAL_ERROR = 3
};

uint32_t gwn_oracle_morton_u32(float nx, float ny, float nz);
uint64_t gwn_oracle_morton_u64(float nx, float ny, float nz);

/* Stable ascending sort of 64-bit keys: out_order[i] = input position of the i-th smallest key. */
void gwn_oracle_stable_order_u64(const uint64_t *keys, uint64_t n, uint32_t *out_order);

int gwn_oracle_build_bvh4_f32(
    const float *vx, const float *vy, const float *vz, uint64_t vertex_count,
    const uint32_t *i0, const uint32_t *i1, const uint32_t *i2, uint64_t triangle_count,
    gwn_oracle_bvh_f32 *out);

/*
 * Build with a chosen binary builder: LBVH (Karras) or H-PLOC with search radius in [1, 8]
 * (gwn_bvh_build_hploc.cuh:104-622).  H-PLOC internal nodes are numbered deterministically:
 * a node whose children hold smallest sorted leaves a < b gets id b - 1.
 */
int gwn_oracle_build_bvh4_ex_f32(
    const float *vx, const float *vy, const float *vz, uint64_t vertex_count,
    const uint32_t *i0, const uint32_t *i1, const uint32_t *i2, uint64_t triangle_count,
    int method, uint32_t hploc_radius, gwn_oracle_bvh_f32 *out);

/*
 * The same for a tree of width 2, 3 or 4 (gwn_bvh_build_collapse.cuh:163-296 with Width): each node
 * expands to at most `width` children; nodes are stored four-wide with the slots past the width
 * empty.  gwn_oracle_build_bvh4_ex_f32 is width 4.
 */
int gwn_oracle_build_bvh_ex_f32(
    const float *vx, const float *vy, const float *vz, uint64_t vertex_count,
    const uint32_t *i0, const uint32_t *i1, const uint32_t *i2, uint64_t triangle_count,
    int method, uint32_t hploc_radius, int width, gwn_oracle_bvh_f32 *out);

/*
 * Bounds refit in place: triangle records from the geometry, child boxes folded in slot order,
 * root bounds (gwn_bvh_refit_impl.cuh:23-192).  Topology, primitive order and stack bounds are
 * untouched.
 */
int gwn_oracle_refit_bvh_f32(
    gwn_oracle_node4_f32 *nodes, uint64_t node_count, gwn_oracle_child_f32 *root,
    const uint32_t *primitive_indices, float *triangles, uint64_t triangle_count,
    const float *vx, const float *vy, const float *vz, uint64_t vertex_count,
    const uint32_t *i0, const uint32_t *i1, const uint32_t *i2);

/* Number of float fields per child slot for Taylor order 0/1/2 (gwn_bvh.cuh:144-202). */
int gwn_oracle_moment_field_count(int order);

/*
 * Moment refit on a canonical BVH4.  Output layout per node is the reference's
 * field-SoA gwn_bvh_moment_node<4, Order, Real>: field f of child slot s is at
 * out[node * 4 * F + f * 4 + s].
 */
int gwn_oracle_refit_moment_f32(
    const gwn_oracle_node4_f32 *nodes, uint64_t node_count, const float *triangles,
    uint64_t triangle_count, int order, float *out);
int gwn_oracle_refit_moment_f64(
    const gwn_oracle_node4_f32 *nodes, uint64_t node_count, const float *triangles,
    uint64_t triangle_count, int order, double *out);

/*
 * Additive payload of one leaf record (gwn_bvh_refit_common.cuh:122-387), 34 values:
 * area, area*P (3), N (3), N_i P_j (9, i-major), raw second moments (18: x,y,z-major of
 * xx yy zz xy yz zx).  Entries beyond the order are zero.
 */
void gwn_oracle_triangle_payload_f64(const float *record, int order, double *out);
void gwn_oracle_triangle_payload_f32(const float *record, int order, float *out);

/* Exact winding over leaf-ordered records, reference-form FP32 arithmetic. */
void gwn_oracle_winding_exact_records_f32(
    const float *triangles, uint64_t triangle_count, const float *qx, const float *qy,
    const float *qz, uint64_t query_count, float *out, int threads);

/* Exact winding in FP64 over the original mesh (exact vertex values). */
void gwn_oracle_winding_exact_mesh_f64(
    const float *vx, const float *vy, const float *vz, const uint32_t *i0, const uint32_t *i1,
    const uint32_t *i2, uint64_t triangle_count, const float *qx, const float *qy,
    const float *qz, uint64_t query_count, double *out, int threads);

/* Signed solid angle of one triangle (a, b, c) seen from q. */
double gwn_oracle_solid_angle_f64(const double *a, const double *b, const double *c,
                                  const double *q);
float gwn_oracle_solid_angle_f32(const float *a, const float *b, const float *c,
                                 const float *q);

/*
 * Taylor traversal replay (FP32, reference-form leaves).  counters (optional,
 * 3 x uint64 per query or NULL): child tests, accepted expansions, leaf
 * triangle evaluations.  Returns NaN for a query whose stack would exceed
 * stack_capacity.  near_fp64 != 0 is the precision split: a leaf triangle whose first
 * vertex is within four edge lengths of the query is evaluated in FP64 on the exact vertices, a
 * farther one in FP32 (unnormalized form); every term, far-field expansions included, is
 * accumulated in FP64.
 */
/* Real = double (no float rounding of the geometry): the exact FP64 sum over a mesh with double
 * vertices, and the double Taylor traversal over double records (9 doubles each, leaf order) and
 * double moments in the field-SoA node layout [node][field][4], on a tree whose child references
 * come from `nodes` (bounds unused). */
int gwn_oracle_refit_moment_d64(const gwn_oracle_node4_f32 *nodes, uint64_t node_count,
                                const double *triangles, uint64_t triangle_count, const double *child_bounds,
                                int order, double *out);
void gwn_oracle_winding_exact_mesh_d64(const double *vx, const double *vy, const double *vz,
                                       const uint32_t *i0, const uint32_t *i1, const uint32_t *i2,
                                       uint64_t triangle_count, const double *qx, const double *qy,
                                       const double *qz, uint64_t query_count, double *out, int threads);
void gwn_oracle_winding_taylor_d64(
    const gwn_oracle_node4_f32 *nodes, uint64_t node_count, uint64_t root_reference,
    const double *triangles, uint64_t triangle_count, const double *moments, int order,
    double accuracy_scale, int stack_capacity, const double *qx, const double *qy, const double *qz,
    uint64_t query_count, double *out, int threads);
void gwn_oracle_winding_taylor_f32(
    const gwn_oracle_node4_f32 *nodes, uint64_t node_count, gwn_oracle_child_f32 root,
    const float *triangles, uint64_t triangle_count, const float *moments, int order,
    float accuracy_scale, int stack_capacity, const float *qx, const float *qy,
    const float *qz, uint64_t query_count, float *out, uint64_t *counters, int near_fp64,
    int threads);

/*
 * The same replay with the analytic gradient (gwn_query_winding_impl.cuh:20-99, 217-325):
 * out (optional) receives the winding number, gx/gy/gz its gradient.
 */
void gwn_oracle_winding_gradient_taylor_f32(
    const gwn_oracle_node4_f32 *nodes, uint64_t node_count, gwn_oracle_child_f32 root,
    const float *triangles, uint64_t triangle_count, const float *moments, int order,
    float accuracy_scale, int stack_capacity, const float *qx, const float *qy,
    const float *qz, uint64_t query_count, float *out, float *gx, float *gy, float *gz,
    uint64_t *counters, int near_fp64, int threads);

/* Exact winding-number gradient in FP64 over the original mesh (sum of triangle gradients). */
void gwn_oracle_winding_gradient_exact_mesh_f64(
    const float *vx, const float *vy, const float *vz, const uint32_t *i0, const uint32_t *i1,
    const uint32_t *i2, uint64_t triangle_count, const float *qx, const float *qy,
    const float *qz, uint64_t query_count, double *gx, double *gy, double *gz, int threads);

/*
 * Unsigned point-to-mesh distance by near-first culled traversal
 * (gwn_query_distance_impl.cuh:36-166): the culling band (clamped at 0) is the initial bound
 * and the result when nothing is closer; NaN when the stack would exceed stack_capacity.
 */
void gwn_oracle_unsigned_distance_f32(
    const gwn_oracle_node4_f32 *nodes, uint64_t node_count, gwn_oracle_child_f32 root,
    const float *triangles, uint64_t triangle_count, const float *qx, const float *qy,
    const float *qz, uint64_t query_count, float culling_band, int stack_capacity, float *out,
    int threads);

/* Ray first-hit record, byte-compatible with gwn_ray_first_hit_result<float, uint32_t>
 * (include/gwn/gwn_query.cuh:74-95): status 0 miss, 1 hit, 2 stack overflow. */
typedef struct {
    float t, u, v, nx, ny, nz;
    uint32_t primitive_id;
    uint8_t status;
    uint8_t pad[3];
} gwn_oracle_ray_hit;

/*
 * Nearest ray-triangle hit on [t_min, t_max] (gwn_query_ray_impl.cuh:20-595) with the batch's
 * traversal dispatch (gwn_query.cuh:276-289): packed near-first when the root is internal, the
 * node count is below 2^29 and packed_stack_bound <= stack_capacity, canonical otherwise
 * (traversal = -1); traversal = 0 / 1 forces the canonical / packed traversal.
 */
void gwn_oracle_ray_first_hit_f32(
    const gwn_oracle_node4_f32 *nodes, uint64_t node_count, gwn_oracle_child_f32 root,
    const float *triangles, const uint32_t *primitive_indices, uint64_t triangle_count,
    uint32_t packed_stack_bound, const float *ox, const float *oy, const float *oz, const float *dx,
    const float *dy, const float *dz, uint64_t ray_count, float t_min, float t_max, int stack_capacity,
    int traversal, gwn_oracle_ray_hit *out, int threads);

/*
 * Algebraic boundary of the oriented triangle chain (gwn_boundary_impl.cuh:21-354): directed rows
 * (start, end, multiplicity) in (min vertex, max vertex) key order.  Output arrays may be NULL
 * to count rows only; otherwise they need room for 3 * triangle_count rows.
 */
int gwn_oracle_boundary_chain(
    const uint32_t *i0, const uint32_t *i1, const uint32_t *i2, uint64_t triangle_count,
    uint32_t *start, uint32_t *end, uint64_t *multiplicity, uint64_t *row_count);

/*
 * Antipodal winding number (out, optional) and its gradient (gx/gy/gz, optional) over the
 * canonical BVH, the original vertices and the boundary chain
 * (gwn_query_winding_antipodal_impl.cuh:55-552).  near_fp64 != 0: boundary areas and area
 * gradients in FP64 on the exact vertices with FP64 sums (precision split).
 */
void gwn_oracle_winding_antipodal_f32(
    const gwn_oracle_node4_f32 *nodes, uint64_t node_count, gwn_oracle_child_f32 root,
    const float *triangles, uint64_t triangle_count, const float *vx, const float *vy, const float *vz,
    const uint32_t *boundary_start, const uint32_t *boundary_end, const uint64_t *boundary_multiplicity,
    uint64_t boundary_edges, const float *qx, const float *qy, const float *qz, uint64_t query_count,
    int stack_capacity, int near_fp64, float *out, float *gx, float *gy, float *gz, int threads);

/*
 * Adjoints (no reference counterpart).  vertices: FP64 (3, V) positions; the BVH, FP32 records and
 * moments give the forward's accept/descend decisions.  L = sum_q u_q w_q (Taylor orders 0-1):
 * the frozen-decision value in FP64, and dL/dV (3, V) and dL/dq (Q, 3).  Distance: brute-force
 * dist, dL/dV and dL/dp for L = sum_p u_p d_p.  Magnitude: per-vertex and per-query sums of the
 * absolute values of every triangle's exact term (error envelopes for FP32 sums).
 */
double gwn_oracle_winding_taylor_frozen_f64(
    const gwn_oracle_node4_f32 *nodes, uint64_t node_count, gwn_oracle_child_f32 root, const float *triangles,
    uint64_t triangle_count, const float *moments, int order, float beta, const uint32_t *primitive_indices,
    const uint32_t *i0, const uint32_t *i1, const uint32_t *i2, const double *vertices, uint64_t vertex_count,
    const float *qx, const float *qy, const float *qz, const double *u, uint64_t query_count, int threads);
int gwn_oracle_winding_taylor_adjoint_f64(
    const gwn_oracle_node4_f32 *nodes, uint64_t node_count, gwn_oracle_child_f32 root, const float *triangles,
    uint64_t triangle_count, const float *moments, int order, float beta, const uint32_t *primitive_indices,
    const uint32_t *i0, const uint32_t *i1, const uint32_t *i2, const double *vertices, uint64_t vertex_count,
    const float *qx, const float *qy, const float *qz, const double *u, uint64_t query_count, double *d_vertices,
    double *d_queries, int threads);
void gwn_oracle_distance_adjoint_f64(
    const double *vertices, uint64_t vertex_count, const uint32_t *i0, const uint32_t *i1, const uint32_t *i2,
    uint64_t triangle_count, const float *px, const float *py, const float *pz, const double *u, uint64_t count,
    double *dist, double *d_vertices, double *d_points, int threads);
void gwn_oracle_adjoint_magnitude_f64(
    const double *vertices, uint64_t vertex_count, const uint32_t *i0, const uint32_t *i1, const uint32_t *i2,
    uint64_t triangle_count, const float *qx, const float *qy, const float *qz, const double *u, uint64_t count,
    double *mag_dv, double *mag_dq, int threads);

/* Squared point-triangle distance (gwn_query_geometry_impl.cuh:46-97). */
float gwn_oracle_point_triangle_distance2_f32(const float *p, const float *a, const float *b, const float *c);

#ifdef __cplusplus
}
#endif

#endif
