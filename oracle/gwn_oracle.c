/*
 * gwn_oracle.c - CPU oracle for the smallgwn geometric-query path.
 *
 * TEST INFRASTRUCTURE ONLY (see gwn_oracle.h).  Plain C11 + pthreads, built
 * with -ffp-contract=off.  Each function cites the reference file:line it
 * restates.
 */
#include "gwn_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define GWN_ORACLE_OFFSET_MASK ((((uint64_t)1) << 47) - 1)
#define GWN_ORACLE_VALID_BIT (((uint64_t)1) << 47)
#define GWN_ORACLE_INVALID_INDEX 0xffffffffu

int gwn_oracle_moment_field_count(int order)
{
    /* gwn_bvh.cuh:144-202: 7, 13 and 23 scalars per child slot. */
    switch (order) {
    case 0: return 7;
    case 1: return 13;
    case 2: return 23;
    default: return -1;
    }
}

#define RECT float
#define REAL float
#define SFX(name) name##f32
#include "gwn_oracle_real.inc"
#undef REAL
#undef SFX

#define REAL double
#define SFX(name) name##f64
#include "gwn_oracle_real.inc"
#undef REAL
#undef SFX
#undef RECT

#define RECT double
#define REAL double
#define SFX(name) name##d64
#include "gwn_oracle_real.inc"
#undef REAL
#undef SFX
#undef RECT

/* ---------------------------------------------------------------------------
 * Morton codes: gwn_bvh_build_common.cuh:35-88
 * ------------------------------------------------------------------------- */

static uint32_t expand_bits_u32(uint32_t value)
{
    uint32_t x = value & 0x000003ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

static uint64_t expand_bits_u64(uint64_t value)
{
    uint64_t x = value & 0x1fffffu;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

static float clamp01(float v)
{
    if (v < 0.0f)
        return 0.0f;
    if (v > 1.0f)
        return 1.0f;
    return v;
}

uint32_t gwn_oracle_morton_u32(float nx, float ny, float nz)
{
    uint32_t x = (uint32_t)(clamp01(nx) * 1023.0f);
    uint32_t y = (uint32_t)(clamp01(ny) * 1023.0f);
    uint32_t z = (uint32_t)(clamp01(nz) * 1023.0f);
    return (expand_bits_u32(x) << 2) | (expand_bits_u32(y) << 1) | expand_bits_u32(z);
}

uint64_t gwn_oracle_morton_u64(float nx, float ny, float nz)
{
    uint64_t x = (uint64_t)(clamp01(nx) * 2097151.0f);
    uint64_t y = (uint64_t)(clamp01(ny) * 2097151.0f);
    uint64_t z = (uint64_t)(clamp01(nz) * 2097151.0f);
    return (expand_bits_u64(x) << 2) | (expand_bits_u64(y) << 1) | expand_bits_u64(z);
}

/* ---------------------------------------------------------------------------
 * Stable LSD radix order (CUB DeviceRadixSort::SortPairs is stable:
 * gwn_bvh_build_common.cuh:325-347).
 * ------------------------------------------------------------------------- */

void gwn_oracle_stable_order_u64(const uint64_t *keys, uint64_t n, uint32_t *out_order)
{
    uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t *src = out_order, *dst = tmp;
    for (uint64_t i = 0; i < n; ++i)
        src[i] = (uint32_t)i;
    for (int shift = 0; shift < 64; shift += 8) {
        uint64_t count[257];
        memset(count, 0, sizeof(count));
        for (uint64_t i = 0; i < n; ++i)
            count[((keys[src[i]] >> shift) & 0xffu) + 1]++;
        for (int d = 0; d < 256; ++d)
            count[d + 1] += count[d];
        for (uint64_t i = 0; i < n; ++i)
            dst[count[(keys[src[i]] >> shift) & 0xffu]++] = src[i];
        uint32_t *t = src;
        src = dst;
        dst = t;
    }
    /* 8 passes: the result is back in out_order. */
    free(tmp);
}

/* ---------------------------------------------------------------------------
 * LBVH (Karras 2012) with the reference's equal-code tie-break:
 * gwn_bvh_build_binary.cuh:37-128
 * ------------------------------------------------------------------------- */

static int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }

static int lbvh_delta(const uint64_t *codes, int64_t n, int64_t i, int64_t j)
{
    if (j < 0 || j >= n)
        return -1;
    uint64_t ci = codes[i], cj = codes[j];
    if (ci == cj) {
        uint64_t diff = (uint64_t)i ^ (uint64_t)j;
        if (diff == 0)
            return 64 + 64;
        return 64 + clz64(diff);
    }
    return clz64(ci ^ cj);
}

static void lbvh_node(const uint64_t *codes, int64_t n, int64_t i, gwn_oracle_binary_node *out,
                      int64_t *first_out)
{
    int dir = (lbvh_delta(codes, n, i, i + 1) - lbvh_delta(codes, n, i, i - 1) >= 0) ? 1 : -1;
    int dmin = lbvh_delta(codes, n, i, i - dir);
    int64_t lmax = 2;
    while (lbvh_delta(codes, n, i, i + lmax * dir) > dmin)
        lmax <<= 1;
    int64_t len = 0;
    for (int64_t step = lmax >> 1; step > 0; step >>= 1)
        if (lbvh_delta(codes, n, i, i + (len + step) * dir) > dmin)
            len += step;
    int64_t j = i + len * dir;
    int64_t first = i < j ? i : j;
    int64_t last = i < j ? j : i;
    int dnode = lbvh_delta(codes, n, first, last);
    int64_t split = first;
    int64_t step = last - first;
    do {
        step = (step + 1) >> 1;
        int64_t cand = split + step;
        if (cand < last && lbvh_delta(codes, n, first, cand) > dnode)
            split = cand;
    } while (step > 1);
    out->left_index = (uint32_t)split;
    out->left_kind = (split == first) ? 2u : 1u;
    out->right_index = (uint32_t)(split + 1);
    out->right_kind = ((split + 1) == last) ? 2u : 1u;
    *first_out = first;
}

/* Half-area used by the collapse expansion rule, gwn_bvh_build_common.cuh:98-104. */
static float half_area(const float *b)
{
    float dx = b[3] - b[0];
    float dy = b[4] - b[1];
    float dz = b[5] - b[2];
    dx = dx < 0.0f ? 0.0f : dx;
    dy = dy < 0.0f ? 0.0f : dy;
    dz = dz < 0.0f ? 0.0f : dz;
    return dx * dy + dy * dz + dz * dx;
}

static void aabb_union(float *dst, const float *a, const float *b)
{
    for (int k = 0; k < 3; ++k) {
        dst[k] = (b[k] < a[k]) ? b[k] : a[k];
        dst[3 + k] = (a[3 + k] < b[3 + k]) ? b[3 + k] : a[3 + k];
    }
}

/* ---------------------------------------------------------------------------
 * H-PLOC binary build, include/gwn/detail/gwn_bvh_build_hploc.cuh:104-478, run sequentially.
 * Every sorted leaf climbs the implicit LBVH of the Morton codes (find_parent_id, :150-156,
 * with the (code, position) delta of :141-149); the first path to reach a boundary publishes its
 * range and stops, the second joins both ranges (:451-470).  A joined range larger than the
 * 16-cluster threshold, or the full range, runs PLOC on the concatenation of its two children's
 * cluster lists (ploc_merge, :420-447): nearest neighbours by the packed key of :195-241 (area
 * bits << 1 with the low four bits replaced by (offset - 1) << 1 and a direction parity), mutual
 * pairs merged with the lower lane keeping the new cluster (:343-347), survivors compacted in
 * lane order (:392-404), until at most 16 clusters remain (1 for the full range).  The list is
 * stored back at the start of the range, 32 lanes wide as on the device (:181-187).
 *
 * Node ids: the reference draws them from an atomic counter, whose order depends on scheduling.
 * Here a merge of clusters whose smallest sorted leaves are a < b creates internal node b - 1,
 * which is unique (a leaf is the smaller-side minimum of at most one merge) and independent of
 * scheduling.  Left child = the lower lane's cluster, as in the reference (:366-367).
 * ------------------------------------------------------------------------- */

enum { HPLOC_LANES = 32, HPLOC_THRESHOLD = 16 };

typedef struct {
    uint64_t T;
    const uint64_t *codes;
    const float *leaf;             /* sorted leaf boxes */
    gwn_oracle_binary_node *nodes;
    float *bounds;
    int64_t *first;
    uint8_t *made;
    uint32_t *clusters;
    uint32_t radius;
    uint64_t merges;
    int failure;
} hploc_state;

static float half_area(const float *b);
static void aabb_union(float *dst, const float *a, const float *b);

static uint64_t hploc_delta(const uint64_t *codes, uint64_t a, uint64_t b)
{
    uint64_t d = codes[a] ^ codes[b];
    return d == 0 ? (a ^ b) : d;
}

static const float *hploc_box(const hploc_state *h, uint32_t c)
{
    return c < h->T ? h->leaf + 6 * (uint64_t)c : h->bounds + 6 * (uint64_t)(c - h->T);
}

static uint32_t hploc_first(const hploc_state *h, uint32_t c)
{
    return c < h->T ? c : (uint32_t)h->first[c - h->T];
}

static uint32_t hploc_area_key(const float *a, const float *b)
{
    float u[6];
    aabb_union(u, a, b);
    float area = half_area(u);
    uint32_t bits;
    memcpy(&bits, &area, sizeof(bits));
    return (bits << 1) & ~15u;
}

static void hploc_merge_range(hploc_state *h, uint64_t left, uint64_t right_end, uint64_t split, int final)
{
    uint32_t idx[HPLOC_LANES];
    float box[HPLOC_LANES][6];
    for (int l = 0; l < HPLOC_LANES; ++l)
        idx[l] = GWN_ORACLE_INVALID_INDEX;
    uint64_t lcl = split - left < HPLOC_THRESHOLD ? split - left : HPLOC_THRESHOLD;
    uint32_t num_left = 0;
    for (uint64_t k = 0; k < lcl; ++k) {
        idx[k] = h->clusters[left + k];
        num_left += idx[k] != GWN_ORACLE_INVALID_INDEX;
    }
    uint64_t lcr = right_end - split < HPLOC_THRESHOLD ? right_end - split : HPLOC_THRESHOLD;
    uint32_t num_right = 0;
    for (uint64_t k = 0; k < lcr; ++k) {
        idx[num_left + k] = h->clusters[split + k];
        num_right += idx[num_left + k] != GWN_ORACLE_INVALID_INDEX;
    }
    uint32_t total = num_left + num_right, num = total;
    for (uint32_t l = 0; l < num; ++l) {
        if (idx[l] == GWN_ORACLE_INVALID_INDEX || idx[l] >= 2 * h->T - 1) {
            h->failure = 1;
            return;
        }
        memcpy(box[l], hploc_box(h, idx[l]), sizeof(float) * 6);
    }
    uint32_t threshold = final ? 1u : (uint32_t)HPLOC_THRESHOLD;
    while (num > threshold) {
        uint32_t nn[HPLOC_LANES];
        for (uint32_t l = 0; l < num; ++l) {
            uint32_t best = 0xffffffffu;
            for (uint32_t r = 1; r <= h->radius; ++r) {
                if (l + r < num) {
                    uint32_t key = hploc_area_key(box[l], box[l + r]) | ((r - 1) << 1) | (l & 1u);
                    best = key < best ? key : best;
                }
                if (l >= r) {
                    uint32_t key = hploc_area_key(box[l - r], box[l]) | ((r - 1) << 1) | ((l & 1u) ^ 1u);
                    best = key < best ? key : best;
                }
            }
            uint32_t off = ((best & 15u) >> 1) + 1u;
            nn[l] = (((best ^ l) & 1u) == 0u) ? l + off : l - off;
        }
        int keep[HPLOC_LANES];
        uint32_t merges = 0;
        for (uint32_t l = 0; l < HPLOC_LANES; ++l) {
            int mutual = 0, merge = 0;
            if (l < num) {
                if (nn[l] >= num || nn[l] == l) {
                    h->failure = 1;
                    return;
                }
                mutual = nn[nn[l]] == l;
                merge = mutual && l < nn[l];
            }
            keep[l] = merge || !mutual;
            if (!merge)
                continue;
            uint32_t a = idx[l], b = idx[nn[l]];
            uint32_t fa = hploc_first(h, a), fb = hploc_first(h, b);
            uint32_t id = (fa > fb ? fa : fb) - 1u;
            if (fa == fb || id >= h->T - 1 || h->made[id]) {
                h->failure = 1;
                return;
            }
            h->made[id] = 1;
            gwn_oracle_binary_node *nd = &h->nodes[id];
            nd->left_index = a < h->T ? a : (uint32_t)(a - h->T);
            nd->left_kind = a < h->T ? 2u : 1u;
            nd->right_index = b < h->T ? b : (uint32_t)(b - h->T);
            nd->right_kind = b < h->T ? 2u : 1u;
            aabb_union(h->bounds + 6 * (uint64_t)id, box[l], box[nn[l]]);
            h->first[id] = fa < fb ? fa : fb;
            idx[l] = (uint32_t)h->T + id;
            memcpy(box[l], h->bounds + 6 * (uint64_t)id, sizeof(float) * 6);
            ++merges;
        }
        if (merges == 0) {
            h->failure = 1;
            return;
        }
        uint32_t k = 0;
        for (uint32_t l = 0; l < HPLOC_LANES; ++l) {
            if (!keep[l])
                continue;
            idx[k] = idx[l];
            memcpy(box[k], box[l], sizeof(float) * 6);
            ++k;
        }
        for (; k < HPLOC_LANES; ++k)
            idx[k] = GWN_ORACLE_INVALID_INDEX;
        num -= merges;
        h->merges += merges;
    }
    for (uint32_t l = 0; l < total; ++l)
        h->clusters[left + l] = idx[l];
}

/* Returns the root cluster id (internal ids offset by T) or INVALID on failure. */
static uint32_t hploc_build(hploc_state *h)
{
    uint64_t T = h->T;
    uint32_t *boundary = (uint32_t *)malloc(sizeof(uint32_t) * T);
    for (uint64_t i = 0; i < T; ++i) {
        h->clusters[i] = (uint32_t)i;
        boundary[i] = GWN_ORACLE_INVALID_INDEX;
    }
    for (uint64_t i = 0; i < T && !h->failure; ++i) {
        uint64_t left = i, right = i, split = 0;
        for (;;) {
            uint32_t prev;
            int to_right = left == 0 ||
                           (right != T - 1 && hploc_delta(h->codes, right, right + 1) <
                                                  hploc_delta(h->codes, left - 1, left));
            if (to_right) {
                prev = boundary[right];
                boundary[right] = (uint32_t)left;
                if (prev != GWN_ORACLE_INVALID_INDEX) {
                    split = right + 1;
                    right = prev;
                }
            } else {
                prev = boundary[left - 1];
                boundary[left - 1] = (uint32_t)right;
                if (prev != GWN_ORACLE_INVALID_INDEX) {
                    split = left;
                    left = prev;
                }
            }
            if (prev == GWN_ORACLE_INVALID_INDEX)
                break;
            uint64_t size = right - left + 1;
            int final = size == T;
            if (size > HPLOC_THRESHOLD || final)
                hploc_merge_range(h, left, right + 1, split, final);
            if (h->failure)
                break;
        }
    }
    free(boundary);
    if (h->failure || h->merges != T - 1)
        return GWN_ORACLE_INVALID_INDEX;
    return h->clusters[0];
}

typedef struct {
    uint32_t binary;     /* binary internal node that roots this wide node */
    uint32_t depth;
    uint32_t first_leaf; /* smallest sorted leaf index under the wide node */
    uint32_t prefix_internal;
    uint32_t prefix_packed;
} wide_task;

int gwn_oracle_build_bvh4_f32(
    const float *vx, const float *vy, const float *vz, uint64_t vertex_count,
    const uint32_t *i0, const uint32_t *i1, const uint32_t *i2, uint64_t triangle_count,
    gwn_oracle_bvh_f32 *out)
{
    return gwn_oracle_build_bvh4_ex_f32(vx, vy, vz, vertex_count, i0, i1, i2, triangle_count,
                                        GWN_ORACLE_BUILD_LBVH, 8u, out);
}

int gwn_oracle_build_bvh4_ex_f32(
    const float *vx, const float *vy, const float *vz, uint64_t vertex_count,
    const uint32_t *i0, const uint32_t *i1, const uint32_t *i2, uint64_t triangle_count,
    int method, uint32_t hploc_radius, gwn_oracle_bvh_f32 *out)
{
    return gwn_oracle_build_bvh_ex_f32(vx, vy, vz, vertex_count, i0, i1, i2, triangle_count, method,
                                       hploc_radius, 4, out);
}

int gwn_oracle_build_bvh_ex_f32(
    const float *vx, const float *vy, const float *vz, uint64_t vertex_count,
    const uint32_t *i0, const uint32_t *i1, const uint32_t *i2, uint64_t triangle_count,
    int method, uint32_t hploc_radius, int width, gwn_oracle_bvh_f32 *out)
{
    uint64_t T = triangle_count;
    out->binary_root = 0;
    if (method != GWN_ORACLE_BUILD_LBVH && method != GWN_ORACLE_BUILD_HPLOC)
        return GWN_ORACLE_INVALID_ARGUMENT;
    if (width < 2 || width > 4)
        return GWN_ORACLE_INVALID_ARGUMENT;
    if (method == GWN_ORACLE_BUILD_HPLOC && (hploc_radius == 0 || hploc_radius > 8))
        return GWN_ORACLE_INVALID_ARGUMENT;
    memset(&out->root, 0, sizeof(out->root));
    out->node_count = 0;
    out->internal_stack_bound = 0;
    out->packed_stack_bound = 0;
    if (T == 0)
        return GWN_ORACLE_OK;
    if (vertex_count == 0 || T > 0x7fffffffull)
        return GWN_ORACLE_INVALID_ARGUMENT;

    /* Scene bounds over every vertex and the host safe inverse (build_common.cuh:243-274). */
    float smin[3] = {vx[0], vy[0], vz[0]}, smax[3] = {vx[0], vy[0], vz[0]};
    for (uint64_t v = 0; v < vertex_count; ++v) {
        float p[3] = {vx[v], vy[v], vz[v]};
        for (int k = 0; k < 3; ++k) {
            if (p[k] < smin[k]) smin[k] = p[k];
            if (p[k] > smax[k]) smax[k] = p[k];
        }
    }
    float inv[3];
    for (int k = 0; k < 3; ++k)
        inv[k] = (smax[k] == smin[k]) ? 1.0f : 1.0f / (smax[k] - smin[k]);
    for (int k = 0; k < 3; ++k) {
        out->scene[k] = smin[k];
        out->scene[3 + k] = smax[k];
        out->scene[6 + k] = inv[k];
    }

    /* Primitive AABBs + Morton codes (build_common.cuh:170-213). */
    float *aabb = (float *)malloc(sizeof(float) * 6 * T);
    uint64_t *codes = (uint64_t *)malloc(sizeof(uint64_t) * T);
    uint32_t *order = out->primitive_indices;
    for (uint64_t t = 0; t < T; ++t) {
        uint32_t a = i0[t], b = i1[t], c = i2[t];
        float *bb = aabb + 6 * t;
        if (a >= vertex_count || b >= vertex_count || c >= vertex_count) {
            memset(bb, 0, sizeof(float) * 6);
            codes[t] = 0;
            continue;
        }
        float X[3] = {vx[a], vx[b], vx[c]}, Y[3] = {vy[a], vy[b], vy[c]}, Z[3] = {vz[a], vz[b], vz[c]};
        const float *P[3] = {X, Y, Z};
        for (int k = 0; k < 3; ++k) {
            const float *q = P[k];
            float lo = (q[2] < q[1]) ? q[2] : q[1];
            lo = (lo < q[0]) ? lo : q[0];
            float hi = (q[1] < q[2]) ? q[2] : q[1];
            hi = (q[0] < hi) ? hi : q[0];
            bb[k] = lo;
            bb[3 + k] = hi;
        }
        float cx = (bb[0] + bb[3]) * 0.5f, cy = (bb[1] + bb[4]) * 0.5f, cz = (bb[2] + bb[5]) * 0.5f;
        codes[t] = gwn_oracle_morton_u64((cx - smin[0]) * inv[0], (cy - smin[1]) * inv[1],
                                         (cz - smin[2]) * inv[2]);
    }
    gwn_oracle_stable_order_u64(codes, T, order);
    float *saabb = (float *)malloc(sizeof(float) * 6 * T);
    for (uint64_t i = 0; i < T; ++i) {
        out->sorted_morton[i] = codes[order[i]];
        memcpy(saabb + 6 * i, aabb + 6 * (uint64_t)order[i], sizeof(float) * 6);
    }

    /* Leaf-ordered records (gwn_bvh_triangle_impl.cuh:23-67, build_impl.cuh:23-41). */
    int status = GWN_ORACLE_OK;
    for (uint64_t i = 0; i < T; ++i) {
        uint32_t p = order[i];
        uint32_t a = i0[p], b = i1[p], c = i2[p];
        if (a >= vertex_count || b >= vertex_count || c >= vertex_count) {
            status = GWN_ORACLE_INTERNAL_ERROR;
            continue;
        }
        float *r = out->triangles + 9 * i;
        r[0] = vx[a]; r[1] = vy[a]; r[2] = vz[a];
        r[3] = vx[b]; r[4] = vy[b]; r[5] = vz[b];
        r[6] = vx[c]; r[7] = vy[c]; r[8] = vz[c];
    }
    if (status != GWN_ORACLE_OK) {
        free(aabb); free(codes); free(saabb);
        return status;
    }

    if (T == 1) {
        memcpy(out->root.bounds, saabb, sizeof(float) * 6);
        out->root.reference = GWN_ORACLE_VALID_BIT | (((uint64_t)1) << 48);
        free(aabb); free(codes); free(saabb);
        return GWN_ORACLE_OK;
    }

    uint64_t NI = T - 1;
    gwn_oracle_binary_node *bn = (gwn_oracle_binary_node *)malloc(sizeof(*bn) * NI);
    int64_t *bfirst = (int64_t *)malloc(sizeof(int64_t) * NI);
    float *bbounds = (float *)malloc(sizeof(float) * 6 * NI);
    uint32_t broot = 0;
    if (method == GWN_ORACLE_BUILD_HPLOC) {
        hploc_state h = {T, out->sorted_morton, saabb, bn, bbounds, bfirst, (uint8_t *)calloc(NI, 1),
                         (uint32_t *)malloc(sizeof(uint32_t) * T), hploc_radius, 0, 0};
        uint32_t root_cluster = hploc_build(&h);
        free(h.made);
        free(h.clusters);
        if (root_cluster == GWN_ORACLE_INVALID_INDEX || root_cluster < T) {
            free(bn); free(bfirst); free(bbounds);
            free(aabb); free(codes); free(saabb);
            return GWN_ORACLE_INTERNAL_ERROR;
        }
        broot = root_cluster - (uint32_t)T;
    } else {
        for (uint64_t i = 0; i < NI; ++i)
            lbvh_node(out->sorted_morton, (int64_t)T, (int64_t)i, &bn[i], &bfirst[i]);
    }
    out->binary_root = broot;

    /* Bottom-up bounds (build_binary.cuh:131-192): post-order over the tree.  H-PLOC already
     * produced every bound as the union of (left, right) at its merge (hploc.cuh:357-359). */
    if (method == GWN_ORACLE_BUILD_LBVH) {
        uint32_t *stack = (uint32_t *)malloc(sizeof(uint32_t) * 2 * (NI + 1));
        uint8_t *visited = (uint8_t *)calloc(NI, 1);
        int64_t sp = 0;
        stack[sp++] = broot;
        while (sp > 0) {
            uint32_t node = stack[sp - 1];
            if (!visited[node]) {
                visited[node] = 1;
                if (bn[node].right_kind == 1u) stack[sp++] = bn[node].right_index;
                if (bn[node].left_kind == 1u) stack[sp++] = bn[node].left_index;
                continue;
            }
            --sp;
            const float *lb = bn[node].left_kind == 1u ? bbounds + 6 * (uint64_t)bn[node].left_index
                                                       : saabb + 6 * (uint64_t)bn[node].left_index;
            const float *rb = bn[node].right_kind == 1u ? bbounds + 6 * (uint64_t)bn[node].right_index
                                                        : saabb + 6 * (uint64_t)bn[node].right_index;
            aabb_union(bbounds + 6 * (uint64_t)node, lb, rb);
        }
        free(stack);
        free(visited);
    }
    if (out->binary_nodes)
        memcpy(out->binary_nodes, bn, sizeof(*bn) * NI);
    if (out->binary_bounds)
        memcpy(out->binary_bounds, bbounds, sizeof(float) * 6 * NI);
    if (out->binary_first)
        for (uint64_t i = 0; i < NI; ++i)
            out->binary_first[i] = (uint32_t)bfirst[i];

    /* BVH4 collapse (collapse.cuh:163-296), breadth-first over a FIFO of tasks.  A node's
     * internal children join the FIFO in ascending order of their smallest sorted leaf, so the
     * FIFO position is the canonical node id: breadth-first, parents in level order, siblings by
     * first leaf.  For LBVH trees (contiguous leaf ranges) that is (depth, first leaf) order. */
    wide_task *tasks = (wide_task *)malloc(sizeof(wide_task) * NI);
    gwn_oracle_node4_f32 *tmp_nodes = (gwn_oracle_node4_f32 *)calloc(NI, sizeof(gwn_oracle_node4_f32));
    uint64_t head = 0, tail = 0;
    tasks[tail++] = (wide_task){broot, 0u, (uint32_t)bfirst[broot], 0u, 0u};
    uint32_t max_int = 0, max_packed = 0;
    while (head < tail) {
        wide_task w = tasks[head];
        uint64_t self = head++;
        gwn_oracle_binary_node root = bn[w.binary];
        uint32_t cidx[4], ckind[4];
        int cnt = 2;
        cidx[0] = root.left_index; ckind[0] = root.left_kind;
        cidx[1] = root.right_index; ckind[1] = root.right_kind;
        while (cnt < width) {  /* collapse.cuh:166-211: expand up to Width candidates */
            int best = -1;
            float best_area = -INFINITY;
            uint32_t best_internal = GWN_ORACLE_INVALID_INDEX;
            for (int s = 0; s < cnt; ++s) {
                if (ckind[s] != 1u)
                    continue;
                float area = half_area(bbounds + 6 * (uint64_t)cidx[s]);
                if (best >= 0 && area < best_area)
                    continue;
                if (best >= 0 && area == best_area && cidx[s] >= best_internal)
                    continue;
                best = s;
                best_area = area;
                best_internal = cidx[s];
            }
            if (best < 0)
                break;
            gwn_oracle_binary_node ex = bn[cidx[best]];
            cidx[best] = ex.left_index; ckind[best] = ex.left_kind;
            cidx[cnt] = ex.right_index; ckind[cnt] = ex.right_kind;
            ++cnt;
        }
        uint32_t n_int = 0;
        for (int s = 0; s < cnt; ++s)
            n_int += (ckind[s] == 1u);
        uint32_t node_int = w.prefix_internal + (n_int ? n_int - 1 : 0);
        uint32_t node_packed = w.prefix_packed + (uint32_t)(cnt - 1);
        if (node_int > max_int) max_int = node_int;
        if (node_packed > max_packed) max_packed = node_packed;
        for (int s = 0; s < cnt; ++s) {
            gwn_oracle_child_f32 *ch = &tmp_nodes[self].children[s];
            if (ckind[s] == 2u) {
                memcpy(ch->bounds, saabb + 6 * (uint64_t)cidx[s], sizeof(float) * 6);
                ch->reference = GWN_ORACLE_VALID_BIT | (uint64_t)cidx[s] | (((uint64_t)1) << 48);
                continue;
            }
            memcpy(ch->bounds, bbounds + 6 * (uint64_t)cidx[s], sizeof(float) * 6);
            /* Rank among this node's internal children by first leaf. */
            uint32_t rank = 0;
            for (int t = 0; t < cnt; ++t)
                rank += ckind[t] == 1u && bfirst[cidx[t]] < bfirst[cidx[s]];
            uint64_t id = tail + rank;
            ch->reference = GWN_ORACLE_VALID_BIT | id;
            tasks[id] = (wide_task){cidx[s], w.depth + 1, (uint32_t)bfirst[cidx[s]], node_int, node_packed};
        }
        tail += n_int;
    }
    uint64_t NW = tail;
    memcpy(out->nodes, tmp_nodes, sizeof(gwn_oracle_node4_f32) * NW);
    memcpy(out->root.bounds, bbounds + 6 * (uint64_t)broot, sizeof(float) * 6);
    out->root.reference = GWN_ORACLE_VALID_BIT;
    out->node_count = NW;
    out->internal_stack_bound = max_int;
    out->packed_stack_bound = max_packed;

    free(tasks); free(tmp_nodes);
    free(bn); free(bfirst); free(bbounds);
    free(aabb); free(codes); free(saabb);
    return GWN_ORACLE_OK;
}

/*
 * Bounds refit (include/gwn/detail/gwn_bvh_refit_impl.cuh:23-127, 129-153, 155-192): rewrite
 * every leaf-ordered triangle record from the geometry through primitive_indices, recompute the
 * leaf boxes exactly as the build does (gwn_bvh_triangle_impl.cuh:56-66), and give every internal
 * child slot the fold of its node's child boxes in slot order, bounds = union(bounds, child)
 * (combine_node, :40-51).  Nodes are numbered breadth-first, so children follow their parents
 * and a reverse sweep visits every child before its parent.  The root of a multi-node tree is
 * the fold over node 0; a leaf root takes its triangle's box (:159-187).
 */
static void triangle_box(const float *r, float *box)
{
    for (int k = 0; k < 3; ++k) {
        float lo = (r[6 + k] < r[3 + k]) ? r[6 + k] : r[3 + k];
        float hi = (r[3 + k] < r[6 + k]) ? r[6 + k] : r[3 + k];
        box[k] = (lo < r[k]) ? lo : r[k];
        box[3 + k] = (r[k] < hi) ? hi : r[k];
    }
}

int gwn_oracle_refit_bvh_f32(
    gwn_oracle_node4_f32 *nodes, uint64_t node_count, gwn_oracle_child_f32 *root,
    const uint32_t *primitive_indices, float *triangles, uint64_t triangle_count,
    const float *vx, const float *vy, const float *vz, uint64_t vertex_count,
    const uint32_t *i0, const uint32_t *i1, const uint32_t *i2)
{
    for (uint64_t s = 0; s < triangle_count; ++s) {
        uint32_t const p = primitive_indices[s];
        if (p >= triangle_count)
            return GWN_ORACLE_INVALID_ARGUMENT;
        uint32_t const v[3] = {i0[p], i1[p], i2[p]};
        for (int j = 0; j < 3; ++j) {
            if (v[j] >= vertex_count)
                return GWN_ORACLE_INVALID_ARGUMENT;
            triangles[9 * s + 3 * j + 0] = vx[v[j]];
            triangles[9 * s + 3 * j + 1] = vy[v[j]];
            triangles[9 * s + 3 * j + 2] = vz[v[j]];
        }
    }
    if (triangle_count == 0)
        return GWN_ORACLE_OK;
    if (node_count == 0) {
        uint64_t const off = root->reference & ((1ull << 47) - 1ull);
        if (off >= triangle_count)
            return GWN_ORACLE_INVALID_ARGUMENT;
        triangle_box(triangles + 9 * off, root->bounds);
        return GWN_ORACLE_OK;
    }
    float *node_box = (float *)malloc(sizeof(float) * 6 * node_count);
    if (!node_box)
        return GWN_ORACLE_INTERNAL_ERROR;
    for (uint64_t n = node_count; n-- > 0;) {
        int init = 0;
        float acc[6];
        for (int s = 0; s < 4; ++s) {
            gwn_oracle_child_f32 *c = &nodes[n].children[s];
            uint64_t const ref = c->reference;
            if (!(ref & (1ull << 47)))
                continue;
            uint64_t const off = ref & ((1ull << 47) - 1ull);
            uint32_t const count = (uint32_t)(ref >> 48);
            float box[6];
            if (count == 0) {
                if (off <= n || off >= node_count) {
                    free(node_box);
                    return GWN_ORACLE_INVALID_ARGUMENT;
                }
                for (int k = 0; k < 6; ++k)
                    box[k] = node_box[6 * off + k];
            } else {
                if (off + count > triangle_count) {
                    free(node_box);
                    return GWN_ORACLE_INVALID_ARGUMENT;
                }
                triangle_box(triangles + 9 * off, box);
                for (uint32_t j = 1; j < count; ++j) {
                    float b2[6];
                    triangle_box(triangles + 9 * (off + j), b2);
                    aabb_union(box, box, b2);
                }
            }
            for (int k = 0; k < 6; ++k)
                c->bounds[k] = box[k];
            if (init) {
                aabb_union(acc, acc, box);
            } else {
                for (int k = 0; k < 6; ++k)
                    acc[k] = box[k];
                init = 1;
            }
        }
        if (!init) {
            free(node_box);
            return GWN_ORACLE_INVALID_ARGUMENT;
        }
        for (int k = 0; k < 6; ++k)
            node_box[6 * n + k] = acc[k];
    }
    for (int k = 0; k < 6; ++k)
        root->bounds[k] = node_box[k];
    free(node_box);
    return GWN_ORACLE_OK;
}

int gwn_oracle_refit_moment_f32(const gwn_oracle_node4_f32 *nodes, uint64_t node_count,
                                const float *triangles, uint64_t triangle_count, int order,
                                float *out)
{
    return refit_moment_impl_f32(nodes, node_count, triangles, triangle_count, order, out, NULL);
}

int gwn_oracle_refit_moment_f64(const gwn_oracle_node4_f32 *nodes, uint64_t node_count,
                                const float *triangles, uint64_t triangle_count, int order,
                                double *out)
{
    return refit_moment_impl_f64(nodes, node_count, triangles, triangle_count, order, out, NULL);
}

/* Real = double moments: double records (leaf order, 9 doubles) and double child boxes
 * (node, slot, 6 = min xyz, max xyz), references from `nodes`. */
int gwn_oracle_refit_moment_d64(const gwn_oracle_node4_f32 *nodes, uint64_t node_count,
                                const double *triangles, uint64_t triangle_count, const double *child_bounds,
                                int order, double *out)
{
    return refit_moment_impl_d64(nodes, node_count, triangles, triangle_count, order, out, child_bounds);
}

void gwn_oracle_triangle_payload_f64(const float *record, int order, double *out)
{
    payload_f64 p;
    triangle_payload_f64(record, order, &p);
    out[0] = p.area; out[1] = p.apx; out[2] = p.apy; out[3] = p.apz;
    out[4] = p.nx; out[5] = p.ny; out[6] = p.nz;
    for (int k = 0; k < 9; ++k) out[7 + k] = p.n1[k];
    for (int k = 0; k < 18; ++k) out[16 + k] = p.raw[k];
}

void gwn_oracle_triangle_payload_f32(const float *record, int order, float *out)
{
    payload_f32 p;
    triangle_payload_f32(record, order, &p);
    out[0] = p.area; out[1] = p.apx; out[2] = p.apy; out[3] = p.apz;
    out[4] = p.nx; out[5] = p.ny; out[6] = p.nz;
    for (int k = 0; k < 9; ++k) out[7 + k] = p.n1[k];
    for (int k = 0; k < 18; ++k) out[16 + k] = p.raw[k];
}

double gwn_oracle_solid_angle_f64(const double *a, const double *b, const double *c,
                                  const double *q)
{
    return solid_angle_f64(a, b, c, q);
}

float gwn_oracle_solid_angle_f32(const float *a, const float *b, const float *c, const float *q)
{
    return solid_angle_f32(a, b, c, q);
}

/* ---------------------------------------------------------------------------
 * Query-parallel helpers
 * ------------------------------------------------------------------------- */

typedef void (*query_range_fn)(void *ctx, uint64_t begin, uint64_t end);

typedef struct {
    query_range_fn fn;
    void *ctx;
    uint64_t begin, end;
} range_job;

static void *range_thread(void *arg)
{
    range_job *job = (range_job *)arg;
    job->fn(job->ctx, job->begin, job->end);
    return NULL;
}

static void parallel_ranges(query_range_fn fn, void *ctx, uint64_t n, int threads)
{
    if (threads < 1)
        threads = 1;
    if ((uint64_t)threads > n)
        threads = n ? (int)n : 1;
    if (threads == 1) {
        fn(ctx, 0, n);
        return;
    }
    pthread_t *tid = (pthread_t *)malloc(sizeof(pthread_t) * threads);
    range_job *jobs = (range_job *)malloc(sizeof(range_job) * threads);
    for (int t = 0; t < threads; ++t) {
        jobs[t].fn = fn;
        jobs[t].ctx = ctx;
        jobs[t].begin = n * (uint64_t)t / (uint64_t)threads;
        jobs[t].end = n * (uint64_t)(t + 1) / (uint64_t)threads;
        pthread_create(&tid[t], NULL, range_thread, &jobs[t]);
    }
    for (int t = 0; t < threads; ++t)
        pthread_join(tid[t], NULL);
    free(tid);
    free(jobs);
}

/* Exact winding over records (gwn_query_winding_impl.cuh:377-389). */
typedef struct {
    const float *tri;
    uint64_t nt;
    const float *qx, *qy, *qz;
    float *out;
} exact_rec_ctx;

static void exact_rec_range(void *vctx, uint64_t begin, uint64_t end)
{
    exact_rec_ctx *c = (exact_rec_ctx *)vctx;
    const float k_pi = (float)3.141592653589793238462643383279502884L;
    for (uint64_t q = begin; q < end; ++q) {
        float qq[3] = {c->qx[q], c->qy[q], c->qz[q]};
        float sum = 0.0f;
        for (uint64_t t = 0; t < c->nt; ++t) {
            float a[3], b[3], cc[3];
            record_vertices_f32(c->tri + 9 * t, a, b, cc);
            sum += solid_angle_f32(a, b, cc, qq);
        }
        c->out[q] = sum / (4.0f * k_pi);
    }
}

void gwn_oracle_winding_exact_records_f32(const float *triangles, uint64_t triangle_count,
                                          const float *qx, const float *qy, const float *qz,
                                          uint64_t query_count, float *out, int threads)
{
    exact_rec_ctx ctx = {triangles, triangle_count, qx, qy, qz, out};
    parallel_ranges(exact_rec_range, &ctx, query_count, threads);
}

/* Exact FP64 winding over the mesh (tests/reference_cpu.cuh:51-82). */
typedef struct {
    const float *vx, *vy, *vz;
    const uint32_t *i0, *i1, *i2;
    uint64_t nt;
    const float *qx, *qy, *qz;
    double *out;
} exact_mesh_ctx;

static void exact_mesh_range(void *vctx, uint64_t begin, uint64_t end)
{
    exact_mesh_ctx *c = (exact_mesh_ctx *)vctx;
    const double k_pi = 3.141592653589793238462643383279502884;
    for (uint64_t q = begin; q < end; ++q) {
        double qq[3] = {c->qx[q], c->qy[q], c->qz[q]};
        double sum = 0.0;
        for (uint64_t t = 0; t < c->nt; ++t) {
            uint32_t ia = c->i0[t], ib = c->i1[t], ic = c->i2[t];
            double a[3] = {c->vx[ia], c->vy[ia], c->vz[ia]};
            double b[3] = {c->vx[ib], c->vy[ib], c->vz[ib]};
            double cc[3] = {c->vx[ic], c->vy[ic], c->vz[ic]};
            sum += solid_angle_f64(a, b, cc, qq);
        }
        c->out[q] = sum / (4.0 * k_pi);
    }
}

void gwn_oracle_winding_exact_mesh_f64(const float *vx, const float *vy, const float *vz,
                                       const uint32_t *i0, const uint32_t *i1, const uint32_t *i2,
                                       uint64_t triangle_count, const float *qx, const float *qy,
                                       const float *qz, uint64_t query_count, double *out,
                                       int threads)
{
    exact_mesh_ctx ctx = {vx, vy, vz, i0, i1, i2, triangle_count, qx, qy, qz, out};
    parallel_ranges(exact_mesh_range, &ctx, query_count, threads);
}

/* ---------------------------------------------------------------------------
 * Taylor traversal replay, gwn_query_winding_impl.cuh:102-374, with or without the gradient
 * (ComputeGradient).
 * ------------------------------------------------------------------------- */

typedef struct {
    const gwn_oracle_node4_f32 *nodes;
    uint64_t node_count;
    gwn_oracle_child_f32 root;
    const float *tri;
    uint64_t nt;
    const float *mom;
    int order, F, cap;
    float beta;
    const float *qx, *qy, *qz;
    float *out;               /* winding (may be NULL) */
    float *gx, *gy, *gz;      /* gradient (NULL: winding only) */
    uint64_t *counters;
    int near_fp64;
} taylor_ctx;

/*
 * Gradient of one triangle's solid angle (Biot-Savart form, gwn_query_winding_impl.cuh:20-61):
 * sum over the edges (i, j) of ((g_i - g_j) . (g_i/|g_i| - g_j/|g_j|)) / |g_i x g_j|^2 (g_i x g_j)
 * with g = vertex - q; zero when q is on a vertex or an edge is seen end-on.
 */
#define GWN_GRAD_TRI(NAME, REALT, SQRT)                                                               \
    static void NAME(const REALT *a, const REALT *b, const REALT *c, const REALT *q, REALT *out)      \
    {                                                                                                 \
        REALT g[3][3], h[3][3], l[3];                                                                 \
        const REALT *v[3] = {a, b, c};                                                                \
        out[0] = out[1] = out[2] = (REALT)0;                                                          \
        for (int i = 0; i < 3; ++i) {                                                                 \
            for (int k = 0; k < 3; ++k)                                                               \
                g[i][k] = v[i][k] - q[k];                                                             \
            l[i] = SQRT(g[i][0] * g[i][0] + g[i][1] * g[i][1] + g[i][2] * g[i][2]);                 \
        }                                                                                             \
        if (l[0] == (REALT)0 || l[1] == (REALT)0 || l[2] == (REALT)0)                                 \
            return;                                                                                   \
        for (int i = 0; i < 3; ++i)                                                                   \
            for (int k = 0; k < 3; ++k)                                                               \
                h[i][k] = g[i][k] / l[i];                                                             \
        for (int e = 0; e < 3; ++e) {                                                                 \
            int i = e, j = (e + 1) % 3;                                                               \
            REALT cr[3] = {g[i][1] * g[j][2] - g[i][2] * g[j][1], g[i][2] * g[j][0] - g[i][0] * g[j][2], \
                           g[i][0] * g[j][1] - g[i][1] * g[j][0]};                                    \
            REALT cs = cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2];                                 \
            REALT t[3] = {(REALT)0, (REALT)0, (REALT)0};                                              \
            if (cs != (REALT)0) {                                                                     \
                REALT d[3] = {g[i][0] - g[j][0], g[i][1] - g[j][1], g[i][2] - g[j][2]};              \
                REALT dh[3] = {h[i][0] - h[j][0], h[i][1] - h[j][1], h[i][2] - h[j][2]};             \
                REALT sc = (d[0] * dh[0] + d[1] * dh[1] + d[2] * dh[2]) / cs;                         \
                t[0] = sc * cr[0]; t[1] = sc * cr[1]; t[2] = sc * cr[2];                              \
            }                                                                                         \
            if (e == 0) {                                                                             \
                out[0] = t[0]; out[1] = t[1]; out[2] = t[2];                                          \
            } else {                                                                                  \
                out[0] += t[0]; out[1] += t[1]; out[2] += t[2];                                       \
            }                                                                                         \
        }                                                                                             \
    }
GWN_GRAD_TRI(grad_tri_f32, float, sqrtf)
GWN_GRAD_TRI(grad_tri_f64, double, sqrt)
#undef GWN_GRAD_TRI

/*
 * Precision split of one leaf triangle (no reference counterpart): within four edge lengths of
 * v0 -- |q - v0|^2 <= 16 max(|e1|^2, |e2|^2) -- the FP64 solid angle (and gradient) of the exact
 * vertices; otherwise FP32 values: the solid angle from the equivalent unnormalized form
 * tan(W/2) = a.(e1 x e2) / (|a||b||c| + (a.b)|c| + (a.c)|b| + (b.c)|a|), the gradient from the
 * exact vertices.  Every value is added to the query's FP64 sums.
 */
static void split_leaf(const float *r, const float *q, const double *qd, double *sum, double *gsum)
{
    float e1[3], e2[3], a[3];
    for (int k = 0; k < 3; ++k) {
        e1[k] = r[3 + k] - r[k];
        e2[k] = r[6 + k] - r[k];
        a[k] = r[k] - q[k];
    }
    float le1 = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
    float le2 = e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2];
    float dq = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    float m = le1 < le2 ? le2 : le1;
    if (!(dq > 16.0f * m)) {
        double a_[3] = {r[0], r[1], r[2]}, b_[3] = {r[3], r[4], r[5]}, c_[3] = {r[6], r[7], r[8]};
        *sum += solid_angle_f64(a_, b_, c_, qd);
        if (gsum) {
            double g[3];
            grad_tri_f64(a_, b_, c_, qd, g);
            gsum[0] += g[0]; gsum[1] += g[1]; gsum[2] += g[2];
        }
        return;
    }
    float b[3] = {a[0] + e1[0], a[1] + e1[1], a[2] + e1[2]};
    float cc[3] = {a[0] + e2[0], a[1] + e2[1], a[2] + e2[2]};
    float la = sqrtf(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    float lb = sqrtf(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    float lc = sqrtf(cc[0] * cc[0] + cc[1] * cc[1] + cc[2] * cc[2]);
    float n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    float num = a[0] * n[0] + a[1] * n[1] + a[2] * n[2];
    float den = la * lb * lc + (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) * lc +
                (a[0] * cc[0] + a[1] * cc[1] + a[2] * cc[2]) * lb + (b[0] * cc[0] + b[1] * cc[1] + b[2] * cc[2]) * la;
    *sum += (double)(2.0f * atan2f(num, den));
    if (gsum) {
        float g[3];
        grad_tri_f32(r, r + 3, r + 6, q, g);
        gsum[0] += (double)g[0]; gsum[1] += (double)g[1]; gsum[2] += (double)g[2];
    }
}

/* One query; writes the winding number and, if grad != NULL, its gradient. */
static float taylor_one(const taylor_ctx *c, float qx, float qy, float qz, uint64_t *cnt, float *grad)
{
    const float k_pi = (float)3.141592653589793238462643383279502884L;
    const double k_pi_d = 3.141592653589793238462643383279502884;
    const float k_inv_4pi = 1.0f / (4.0f * k_pi);
    float q[3] = {qx, qy, qz};
    double qd[3] = {qx, qy, qz};
    float omega = 0.0f;
    double near = 0.0;
    float gs[3] = {0.0f, 0.0f, 0.0f};
    double gd[3] = {0.0, 0.0, 0.0};
    float beta2 = c->beta * c->beta;
    uint32_t stack[1024];
    int sp = 0;
    uint64_t tests = 0, accepted = 0, leaves = 0;
#define LEAF(child_ptr)                                                                            \
    do {                                                                                           \
        uint64_t off_ = (child_ptr)->reference & GWN_ORACLE_OFFSET_MASK;                          \
        uint32_t n_ = (uint32_t)((child_ptr)->reference >> 48);                                   \
        for (uint32_t k_ = 0; k_ < n_; ++k_) {                                                    \
            if (off_ + k_ >= c->nt)                                                               \
                continue;                                                                         \
            const float *r_ = c->tri + 9 * (off_ + k_);                                           \
            if (c->near_fp64) {                                                                   \
                split_leaf(r_, q, qd, &near, grad ? gd : NULL);                                   \
            } else {                                                                              \
                float a_[3], b_[3], c_[3];                                                        \
                record_vertices_f32(r_, a_, b_, c_);                                              \
                omega += solid_angle_f32(a_, b_, c_, q);                                          \
                if (grad) {                                                                       \
                    float g_[3];                                                                  \
                    grad_tri_f32(a_, b_, c_, q, g_);                                              \
                    gs[0] += g_[0]; gs[1] += g_[1]; gs[2] += g_[2];                               \
                }                                                                                 \
            }                                                                                     \
            ++leaves;                                                                             \
        }                                                                                         \
    } while (0)
#define FINISH()                                                                                   \
    do {                                                                                           \
        if (cnt) { cnt[0] = tests; cnt[1] = accepted; cnt[2] = leaves; }                          \
        if (grad) {                                                                                \
            for (int k_ = 0; k_ < 3; ++k_)                                                         \
                grad[k_] = c->near_fp64 ? (float)(gd[k_] * (1.0 / (4.0 * k_pi_d))) : gs[k_] * k_inv_4pi; \
        }                                                                                          \
        return c->near_fp64 ? (float)(((double)omega + near) / (4.0 * k_pi_d)) : omega / (4.0f * k_pi); \
    } while (0)
    if ((c->root.reference & GWN_ORACLE_VALID_BIT) && (c->root.reference >> 48) != 0) {
        LEAF(&c->root);
        FINISH();
    }
    uint32_t node = (uint32_t)(c->root.reference & GWN_ORACLE_OFFSET_MASK);
    for (;;) {
        if (node >= c->node_count)
            break;
        const gwn_oracle_node4_f32 *nd = &c->nodes[node];
        const float *t = c->mom + (size_t)node * 4 * c->F;
        uint32_t next = GWN_ORACLE_INVALID_INDEX;
        for (int s = 0; s < 4; ++s) {
            const gwn_oracle_child_f32 *ch = &nd->children[s];
            if (!(ch->reference & GWN_ORACLE_VALID_BIT))
                continue;
            ++tests;
#define T(f) t[(f) * 4 + s]
            float qrx = qx - T(1), qry = qy - T(2), qrz = qz - T(3);
            float ql2 = qrx * qrx + qry * qry + qrz * qrz;
            int descend = !(ql2 > 0.0f);
            if (!descend)
                descend = ql2 <= T(0) * beta2;
            if (!descend) {
                float m1 = 1.0f / sqrtf(ql2);
                float m2 = m1 * m1;
                float m3 = m2 * m1;
                float qnx = qrx * m1, qny = qry * m1, qnz = qrz * m1;
                float Nx = T(4), Ny = T(5), Nz = T(6);
                float rdotN = qnx * Nx + qny * Ny + qnz * Nz;
                float w = -m2 * rdotN;
                float g[3] = {0.0f, 0.0f, 0.0f};
                if (grad) {
                    g[0] = m3 * (3.0f * qnx * rdotN - Nx);
                    g[1] = m3 * (3.0f * qny * rdotN - Ny);
                    g[2] = m3 * (3.0f * qnz * rdotN - Nz);
                }
                if (c->order >= 1) {
                    float qxx = qnx * qnx, qyy = qny * qny, qzz = qnz * qnz;
                    float Txx = T(7), Tyy = T(8), Tzz = T(9);
                    float Sxy = T(10), Syz = T(11), Szx = T(12);
                    float Tr = Txx + Tyy + Tzz;
                    float S = qxx * Txx + qyy * Tyy + qzz * Tzz + qnx * qny * Sxy +
                              qnx * qnz * Szx + qny * qnz * Syz;
                    w += m3 * (Tr - 3.0f * S);
                    if (grad) {
                        /* gwn_query_winding_impl.cuh:247-257 */
                        float m4 = m2 * m2;
                        float vx = 2.0f * Txx * qnx + Sxy * qny + Szx * qnz;
                        float vy = Sxy * qnx + 2.0f * Tyy * qny + Syz * qnz;
                        float vz = Szx * qnx + Syz * qny + 2.0f * Tzz * qnz;
                        float coeff = 15.0f * S - 3.0f * Tr;
                        g[0] += m4 * (coeff * qnx - 3.0f * vx);
                        g[1] += m4 * (coeff * qny - 3.0f * vy);
                        g[2] += m4 * (coeff * qnz - 3.0f * vz);
                    }
                }
                if (c->order >= 2) {
                    float m4 = m2 * m2;
                    float qnx2 = qnx * qnx, qny2 = qny * qny, qnz2 = qnz * qnz;
                    float qnx3 = qnx2 * qnx, qny3 = qny2 * qny, qnz3 = qnz2 * qnz;
                    float xxx = T(13), yyy = T(14), zzz = T(15), spn = T(16);
                    float Axy = T(17), Axz = T(18), Ayz = T(19), Ayx = T(20), Azx = T(21), Azy = T(22);
                    float t0x = Ayx + Azx, t0y = Azy + Axy, t0z = Axz + Ayz;
                    float t1x = qny * Axy + qnz * Axz;
                    float t1y = qnz * Ayz + qnx * Ayx;
                    float t1z = qnx * Azx + qny * Azy;
                    w += m4 * (1.5f * (qnx * (3.0f * xxx + t0x) + qny * (3.0f * yyy + t0y) +
                                       qnz * (3.0f * zzz + t0z)) -
                               7.5f * (qnx3 * xxx + qny3 * yyy + qnz3 * zzz + qnx * qny * qnz * spn +
                                       qnx2 * t1x + qny2 * t1y + qnz2 * t1z));
                    if (grad) {
                        /* gwn_query_winding_impl.cuh:295-323 */
                        float m5 = m4 * m1;
                        float Px = 3.0f * xxx + Ayx + Azx;
                        float Py = 3.0f * yyy + Azy + Axy;
                        float Pz = 3.0f * zzz + Axz + Ayz;
                        float L2 = qnx * Px + qny * Py + qnz * Pz;
                        float C2 = qnx3 * xxx + qny3 * yyy + qnz3 * zzz + qnx * qny * qnz * spn +
                                   qnx2 * (qny * Axy + qnz * Axz) + qny2 * (qnz * Ayz + qnx * Ayx) +
                                   qnz2 * (qnx * Azx + qny * Azy);
                        float Rx = 3.0f * xxx * qnx2 + 2.0f * Axy * qnx * qny + 2.0f * Axz * qnx * qnz +
                                   Ayx * qny2 + Azx * qnz2 + spn * qny * qnz;
                        float Ry = 3.0f * yyy * qny2 + 2.0f * Ayx * qnx * qny + 2.0f * Ayz * qny * qnz +
                                   Axy * qnx2 + Azy * qnz2 + spn * qnx * qnz;
                        float Rz = 3.0f * zzz * qnz2 + 2.0f * Azx * qnx * qnz + 2.0f * Azy * qny * qnz +
                                   Axz * qnx2 + Ayz * qny2 + spn * qnx * qny;
                        float h = 52.5f * C2;
                        g[0] += m5 * (1.5f * Px - 7.5f * (qnx * L2 + Rx) + qnx * h);
                        g[1] += m5 * (1.5f * Py - 7.5f * (qny * L2 + Ry) + qny * h);
                        g[2] += m5 * (1.5f * Pz - 7.5f * (qnz * L2 + Rz) + qnz * h);
                    }
                }
#undef T
                if (isfinite(w)) {
                    if (c->near_fp64) {
                        near += (double)w; /* the split accumulates every term in FP64 */
                        gd[0] += (double)g[0]; gd[1] += (double)g[1]; gd[2] += (double)g[2];
                    } else {
                        omega += w;
                        gs[0] += g[0]; gs[1] += g[1]; gs[2] += g[2];
                    }
                    ++accepted;
                    continue;
                }
            }
            uint64_t ref = ch->reference;
            if ((ref >> 48) == 0) {
                if (next != GWN_ORACLE_INVALID_INDEX) {
                    if (sp >= c->cap || sp >= 1024) {
                        if (cnt) { cnt[0] = tests; cnt[1] = accepted; cnt[2] = leaves; }
                        if (grad)
                            grad[0] = grad[1] = grad[2] = NAN;
                        return NAN;
                    }
                    stack[sp++] = next;
                }
                next = (uint32_t)(ref & GWN_ORACLE_OFFSET_MASK);
                continue;
            }
            LEAF(ch);
        }
        if (next != GWN_ORACLE_INVALID_INDEX) {
            node = next;
            continue;
        }
        if (sp == 0)
            break;
        node = stack[--sp];
    }
#undef LEAF
    FINISH();
#undef FINISH
}

static void taylor_range(void *vctx, uint64_t begin, uint64_t end)
{
    taylor_ctx *c = (taylor_ctx *)vctx;
    for (uint64_t q = begin; q < end; ++q) {
        float g[3];
        float w = taylor_one(c, c->qx[q], c->qy[q], c->qz[q], c->counters ? c->counters + 3 * q : NULL,
                             c->gx ? g : NULL);
        if (c->out)
            c->out[q] = w;
        if (c->gx) {
            c->gx[q] = g[0];
            c->gy[q] = g[1];
            c->gz[q] = g[2];
        }
    }
}

void gwn_oracle_winding_taylor_f32(
    const gwn_oracle_node4_f32 *nodes, uint64_t node_count, gwn_oracle_child_f32 root,
    const float *triangles, uint64_t triangle_count, const float *moments, int order,
    float accuracy_scale, int stack_capacity, const float *qx, const float *qy,
    const float *qz, uint64_t query_count, float *out, uint64_t *counters, int near_fp64,
    int threads)
{
    taylor_ctx ctx = {nodes, node_count, root, triangles, triangle_count, moments, order,
                      gwn_oracle_moment_field_count(order), stack_capacity, accuracy_scale,
                      qx, qy, qz, out, NULL, NULL, NULL, counters, near_fp64};
    parallel_ranges(taylor_range, &ctx, query_count, threads);
}

/* ---------------------------------------------------------------------------
 * Real = double: the exact sum over a mesh with double vertices, and the Taylor traversal
 * (gwn_query_winding_impl.cuh:102-374 instantiated with double) over double records and double
 * moments on a tree whose references come from the float build.  Records are (v0, v1, v2) as 9
 * doubles; the corners are v0, v0 + (v1 - v0), v0 + (v2 - v0) as the reference rebuilds them.
 * ------------------------------------------------------------------------- */

typedef struct {
    const double *vx, *vy, *vz;
    const uint32_t *i0, *i1, *i2;
    uint64_t nt;
    const double *qx, *qy, *qz;
    double *out;
} exact_mesh_d_ctx;

static void exact_mesh_d_range(void *vctx, uint64_t begin, uint64_t end)
{
    exact_mesh_d_ctx *c = (exact_mesh_d_ctx *)vctx;
    const double k_pi = 3.141592653589793238462643383279502884;
    for (uint64_t q = begin; q < end; ++q) {
        double qq[3] = {c->qx[q], c->qy[q], c->qz[q]};
        double sum = 0.0;
        for (uint64_t t = 0; t < c->nt; ++t) {
            uint32_t ia = c->i0[t], ib = c->i1[t], ic = c->i2[t];
            double a[3] = {c->vx[ia], c->vy[ia], c->vz[ia]};
            double b[3] = {c->vx[ib], c->vy[ib], c->vz[ib]};
            double cc[3] = {c->vx[ic], c->vy[ic], c->vz[ic]};
            sum += solid_angle_f64(a, b, cc, qq);
        }
        c->out[q] = sum / (4.0 * k_pi);
    }
}

void gwn_oracle_winding_exact_mesh_d64(const double *vx, const double *vy, const double *vz,
                                       const uint32_t *i0, const uint32_t *i1, const uint32_t *i2,
                                       uint64_t triangle_count, const double *qx, const double *qy,
                                       const double *qz, uint64_t query_count, double *out, int threads)
{
    exact_mesh_d_ctx ctx = {vx, vy, vz, i0, i1, i2, triangle_count, qx, qy, qz, out};
    parallel_ranges(exact_mesh_d_range, &ctx, query_count, threads);
}

static void record_vertices_d(const double *r, double *a, double *b, double *c)
{
    for (int k = 0; k < 3; ++k) {
        double e1 = r[3 + k] - r[k], e2 = r[6 + k] - r[k];
        a[k] = r[k];
        b[k] = r[k] + e1;
        c[k] = r[k] + e2;
    }
}

typedef struct {
    const gwn_oracle_node4_f32 *nodes;
    uint64_t node_count;
    uint64_t root_reference;
    const double *tri;
    uint64_t nt;
    const double *mom;
    int order, F, cap;
    double beta;
    const double *qx, *qy, *qz;
    double *out;
} taylor_d_ctx;

static double taylor_leaf_d(const taylor_d_ctx *c, uint64_t ref, const double *q)
{
    uint64_t off = ref & GWN_ORACLE_OFFSET_MASK;
    uint32_t n = (uint32_t)(ref >> 48);
    double sum = 0.0;
    for (uint32_t k = 0; k < n; ++k) {
        if (off + k >= c->nt)
            continue;
        double a[3], b[3], cc[3];
        record_vertices_d(c->tri + 9 * (off + k), a, b, cc);
        sum += solid_angle_f64(a, b, cc, q);
    }
    return sum;
}

static double taylor_one_d(const taylor_d_ctx *c, double qx, double qy, double qz)
{
    const double k_pi = 3.141592653589793238462643383279502884;
    double q[3] = {qx, qy, qz};
    double omega = 0.0;
    double beta2 = c->beta * c->beta;
    uint32_t stack[1024];
    int sp = 0;
    if ((c->root_reference & GWN_ORACLE_VALID_BIT) && (c->root_reference >> 48) != 0)
        return taylor_leaf_d(c, c->root_reference, q) / (4.0 * k_pi);
    uint32_t node = (uint32_t)(c->root_reference & GWN_ORACLE_OFFSET_MASK);
    for (;;) {
        if (node >= c->node_count)
            break;
        const gwn_oracle_node4_f32 *nd = &c->nodes[node];
        const double *t = c->mom + (size_t)node * 4 * c->F;
        uint32_t next = GWN_ORACLE_INVALID_INDEX;
        for (int s = 0; s < 4; ++s) {
            uint64_t ref = nd->children[s].reference;
            if (!(ref & GWN_ORACLE_VALID_BIT))
                continue;
#define T(f) t[(f) * 4 + s]
            double qrx = qx - T(1), qry = qy - T(2), qrz = qz - T(3);
            double ql2 = qrx * qrx + qry * qry + qrz * qrz;
            int descend = !(ql2 > 0.0) || ql2 <= T(0) * beta2;
            if (!descend) {
                double m1 = 1.0 / sqrt(ql2);
                double m2 = m1 * m1, m3 = m2 * m1;
                double qnx = qrx * m1, qny = qry * m1, qnz = qrz * m1;
                double w = -m2 * (qnx * T(4) + qny * T(5) + qnz * T(6));
                if (c->order >= 1) {
                    double Txx = T(7), Tyy = T(8), Tzz = T(9);
                    double S = qnx * qnx * Txx + qny * qny * Tyy + qnz * qnz * Tzz + qnx * qny * T(10) +
                               qnx * qnz * T(12) + qny * qnz * T(11);
                    w += m3 * ((Txx + Tyy + Tzz) - 3.0 * S);
                }
                if (c->order >= 2) {
                    double m4 = m2 * m2;
                    double qnx2 = qnx * qnx, qny2 = qny * qny, qnz2 = qnz * qnz;
                    double xxx = T(13), yyy = T(14), zzz = T(15), spn = T(16);
                    double Axy = T(17), Axz = T(18), Ayz = T(19), Ayx = T(20), Azx = T(21), Azy = T(22);
                    double t1x = qny * Axy + qnz * Axz, t1y = qnz * Ayz + qnx * Ayx, t1z = qnx * Azx + qny * Azy;
                    w += m4 * (1.5 * (qnx * (3.0 * xxx + Ayx + Azx) + qny * (3.0 * yyy + Azy + Axy) +
                                      qnz * (3.0 * zzz + Axz + Ayz)) -
                               7.5 * (qnx2 * qnx * xxx + qny2 * qny * yyy + qnz2 * qnz * zzz +
                                      qnx * qny * qnz * spn + qnx2 * t1x + qny2 * t1y + qnz2 * t1z));
                }
#undef T
                if (isfinite(w)) {
                    omega += w;
                    continue;
                }
            }
            if ((ref >> 48) == 0) {
                if (next != GWN_ORACLE_INVALID_INDEX) {
                    if (sp >= c->cap || sp >= 1024)
                        return NAN;
                    stack[sp++] = next;
                }
                next = (uint32_t)(ref & GWN_ORACLE_OFFSET_MASK);
                continue;
            }
            omega += taylor_leaf_d(c, ref, q);
        }
        if (next != GWN_ORACLE_INVALID_INDEX) {
            node = next;
            continue;
        }
        if (sp == 0)
            break;
        node = stack[--sp];
    }
    return omega / (4.0 * k_pi);
}

static void taylor_d_range(void *vctx, uint64_t begin, uint64_t end)
{
    taylor_d_ctx *c = (taylor_d_ctx *)vctx;
    for (uint64_t q = begin; q < end; ++q)
        c->out[q] = taylor_one_d(c, c->qx[q], c->qy[q], c->qz[q]);
}

void gwn_oracle_winding_taylor_d64(
    const gwn_oracle_node4_f32 *nodes, uint64_t node_count, uint64_t root_reference,
    const double *triangles, uint64_t triangle_count, const double *moments, int order,
    double accuracy_scale, int stack_capacity, const double *qx, const double *qy, const double *qz,
    uint64_t query_count, double *out, int threads)
{
    taylor_d_ctx ctx = {nodes, node_count, root_reference, triangles, triangle_count, moments, order,
                        gwn_oracle_moment_field_count(order), stack_capacity, accuracy_scale, qx, qy, qz, out};
    parallel_ranges(taylor_d_range, &ctx, query_count, threads);
}

void gwn_oracle_winding_gradient_taylor_f32(
    const gwn_oracle_node4_f32 *nodes, uint64_t node_count, gwn_oracle_child_f32 root,
    const float *triangles, uint64_t triangle_count, const float *moments, int order,
    float accuracy_scale, int stack_capacity, const float *qx, const float *qy,
    const float *qz, uint64_t query_count, float *out, float *gx, float *gy, float *gz,
    uint64_t *counters, int near_fp64, int threads)
{
    taylor_ctx ctx = {nodes, node_count, root, triangles, triangle_count, moments, order,
                      gwn_oracle_moment_field_count(order), stack_capacity, accuracy_scale,
                      qx, qy, qz, out, gx, gy, gz, counters, near_fp64};
    parallel_ranges(taylor_range, &ctx, query_count, threads);
}

/* Exact gradient of the winding number over all triangles in FP64 (exact vertices). */
typedef struct {
    const float *vx, *vy, *vz;
    const uint32_t *i0, *i1, *i2;
    uint64_t nt;
    const float *qx, *qy, *qz;
    double *gx, *gy, *gz;
} grad_mesh_ctx;

static void grad_mesh_range(void *vctx, uint64_t begin, uint64_t end)
{
    grad_mesh_ctx *c = (grad_mesh_ctx *)vctx;
    const double k_pi_d = 3.141592653589793238462643383279502884;
    for (uint64_t q = begin; q < end; ++q) {
        double qq[3] = {c->qx[q], c->qy[q], c->qz[q]};
        double s[3] = {0.0, 0.0, 0.0};
        for (uint64_t t = 0; t < c->nt; ++t) {
            double a[3] = {c->vx[c->i0[t]], c->vy[c->i0[t]], c->vz[c->i0[t]]};
            double b[3] = {c->vx[c->i1[t]], c->vy[c->i1[t]], c->vz[c->i1[t]]};
            double cc[3] = {c->vx[c->i2[t]], c->vy[c->i2[t]], c->vz[c->i2[t]]};
            double g[3];
            grad_tri_f64(a, b, cc, qq, g);
            s[0] += g[0]; s[1] += g[1]; s[2] += g[2];
        }
        c->gx[q] = s[0] / (4.0 * k_pi_d);
        c->gy[q] = s[1] / (4.0 * k_pi_d);
        c->gz[q] = s[2] / (4.0 * k_pi_d);
    }
}

void gwn_oracle_winding_gradient_exact_mesh_f64(
    const float *vx, const float *vy, const float *vz, const uint32_t *i0, const uint32_t *i1,
    const uint32_t *i2, uint64_t triangle_count, const float *qx, const float *qy,
    const float *qz, uint64_t query_count, double *gx, double *gy, double *gz, int threads)
{
    grad_mesh_ctx ctx = {vx, vy, vz, i0, i1, i2, triangle_count, qx, qy, qz, gx, gy, gz};
    parallel_ranges(grad_mesh_range, &ctx, query_count, threads);
}

/* ---------------------------------------------------------------------------
 * Unsigned distance, include/gwn/detail/gwn_query_distance_impl.cuh:20-192 and
 * gwn_query_geometry_impl.cuh:46-114.
 * ------------------------------------------------------------------------- */

static float dot3f(const float *a, const float *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* Squared distance from p to the closed triangle (a, b, c): Voronoi regions (Ericson). */
static float point_triangle_distance2_f32(const float *p, const float *a, const float *b, const float *c)
{
    float ab[3], ac[3], ap[3], bp[3], cp[3], r[3];
    for (int k = 0; k < 3; ++k) {
        ab[k] = b[k] - a[k];
        ac[k] = c[k] - a[k];
        ap[k] = p[k] - a[k];
    }
    float d1 = dot3f(ab, ap), d2 = dot3f(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f)
        return dot3f(ap, ap);
    for (int k = 0; k < 3; ++k)
        bp[k] = p[k] - b[k];
    float d3 = dot3f(ab, bp), d4 = dot3f(ac, bp);
    if (d3 >= 0.0f && d4 <= d3)
        return dot3f(bp, bp);
    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        float v = d1 / (d1 - d3);
        for (int k = 0; k < 3; ++k)
            r[k] = p[k] - (a[k] + v * ab[k]);
        return dot3f(r, r);
    }
    for (int k = 0; k < 3; ++k)
        cp[k] = p[k] - c[k];
    float d5 = dot3f(ab, cp), d6 = dot3f(ac, cp);
    if (d6 >= 0.0f && d5 <= d6)
        return dot3f(cp, cp);
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        float w = d2 / (d2 - d6);
        for (int k = 0; k < 3; ++k)
            r[k] = p[k] - (a[k] + w * ac[k]);
        return dot3f(r, r);
    }
    float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        for (int k = 0; k < 3; ++k)
            r[k] = p[k] - (b[k] + w * (c[k] - b[k]));
        return dot3f(r, r);
    }
    float denom = 1.0f / (va + vb + vc);
    float v = vb * denom, w = vc * denom;
    for (int k = 0; k < 3; ++k)
        r[k] = p[k] - (a[k] + v * ab[k] + w * ac[k]);
    return dot3f(r, r);
}

static float fmax0(float x) { return (x < 0.0f) ? 0.0f : x; } /* std::max(x, 0) */

static float aabb_distance2_f32(const float *q, const float *b)
{
    float dx = fmax0(b[0] - q[0]) + fmax0(q[0] - b[3]);
    float dy = fmax0(b[1] - q[1]) + fmax0(q[1] - b[4]);
    float dz = fmax0(b[2] - q[2]) + fmax0(q[2] - b[5]);
    return dx * dx + dy * dy + dz * dz;
}

typedef struct {
    const gwn_oracle_node4_f32 *nodes;
    uint64_t node_count;
    gwn_oracle_child_f32 root;
    const float *tri;
    uint64_t nt;
    const float *qx, *qy, *qz;
    float band;
    int cap;
    float *out;
} distance_ctx;

/* Returns 0 when a leaf reaches distance 0 (the global lower bound), like visit_leaf. */
static int distance_leaf(const distance_ctx *c, uint64_t ref, const float *q, float *best)
{
    uint64_t off = ref & GWN_ORACLE_OFFSET_MASK;
    uint32_t n = (uint32_t)(ref >> 48);
    for (uint32_t k = 0; k < n; ++k) {
        if (off + k >= c->nt)
            continue;
        float a[3], b[3], cc[3];
        record_vertices_f32(c->tri + 9 * (off + k), a, b, cc);
        float d2 = point_triangle_distance2_f32(q, a, b, cc);
        if (d2 < *best)
            *best = d2;
    }
    return *best > 0.0f;
}

static float distance_one(const distance_ctx *c, float qx, float qy, float qz)
{
    float q[3] = {qx, qy, qz};
    float band = c->band < 0.0f ? 0.0f : c->band; /* std::max(culling_band, 0) */
    float best = band * band;
    uint64_t stack[1024];
    int sp = 0;
    uint64_t reference = c->root.reference;
    for (;;) {
        uint64_t next_ref = 0;
        int valid = (reference & GWN_ORACLE_VALID_BIT) != 0;
        uint32_t count = (uint32_t)(reference >> 48);
        uint64_t off = reference & GWN_ORACLE_OFFSET_MASK;
        if (valid && count != 0) {
            if (!distance_leaf(c, reference, q, &best))
                return 0.0f;
        } else if (valid && count == 0 && off < c->node_count) {
            const gwn_oracle_node4_f32 *nd = &c->nodes[off];
            float cd[4];
            uint64_t cr[4];
            int m = 0;
            for (int s = 0; s < 4; ++s) {
                const gwn_oracle_child_f32 *ch = &nd->children[s];
                if (!(ch->reference & GWN_ORACLE_VALID_BIT))
                    continue;
                float d2 = aabb_distance2_f32(q, ch->bounds);
                if (d2 >= best)
                    continue;
                cd[m] = d2;
                cr[m] = ch->reference;
                ++m;
            }
            for (int s = m; s < 4; ++s) {
                cd[s] = INFINITY;
                cr[s] = 0;
            }
#define CSWAP(i, j)                                                                                \
            if (cd[i] > cd[j]) {                                                                   \
                float td = cd[i]; cd[i] = cd[j]; cd[j] = td;                                       \
                uint64_t tr = cr[i]; cr[i] = cr[j]; cr[j] = tr;                                    \
            }
            CSWAP(0, 1) CSWAP(2, 3) CSWAP(0, 2) CSWAP(1, 3) CSWAP(1, 2)
#undef CSWAP
            for (int s = 0; s < m; ++s) {
                if ((cr[s] & GWN_ORACLE_VALID_BIT) && (cr[s] >> 48) != 0 && cd[s] < best &&
                    !distance_leaf(c, cr[s], q, &best))
                    return 0.0f;
            }
            for (int s = m - 1; s >= 0; --s) {
                if (!((cr[s] & GWN_ORACLE_VALID_BIT) && (cr[s] >> 48) == 0) || cd[s] >= best)
                    continue;
                if (next_ref != 0) {
                    if (sp >= c->cap || sp >= 1024)
                        return NAN;
                    stack[sp++] = next_ref;
                }
                next_ref = cr[s];
            }
        }
        if (next_ref != 0) {
            reference = next_ref;
            continue;
        }
        if (sp == 0)
            break;
        reference = stack[--sp];
    }
    return sqrtf(best);
}

static void distance_range(void *vctx, uint64_t begin, uint64_t end)
{
    distance_ctx *c = (distance_ctx *)vctx;
    for (uint64_t q = begin; q < end; ++q)
        c->out[q] = distance_one(c, c->qx[q], c->qy[q], c->qz[q]);
}

void gwn_oracle_unsigned_distance_f32(
    const gwn_oracle_node4_f32 *nodes, uint64_t node_count, gwn_oracle_child_f32 root,
    const float *triangles, uint64_t triangle_count, const float *qx, const float *qy,
    const float *qz, uint64_t query_count, float culling_band, int stack_capacity, float *out,
    int threads)
{
    distance_ctx ctx = {nodes, node_count, root, triangles, triangle_count, qx, qy, qz,
                        culling_band, stack_capacity, out};
    parallel_ranges(distance_range, &ctx, query_count, threads);
}

float gwn_oracle_point_triangle_distance2_f32(const float *p, const float *a, const float *b, const float *c)
{
    return point_triangle_distance2_f32(p, a, b, c);
}

/* ---------------------------------------------------------------------------
 * Ray first hit, include/gwn/detail/gwn_query_ray_impl.cuh:20-595 and the batch dispatch of
 * include/gwn/gwn_query.cuh:214-290 (packed near-first traversal when the root is internal, the
 * node count is below 2^29 and packed_stack_bound fits the stack; canonical traversal otherwise).
 * ------------------------------------------------------------------------- */

static float fmin_std(float a, float b) { return (b < a) ? b : a; } /* std::min */
static float fmax_std(float a, float b) { return (a < b) ? b : a; } /* std::max */

typedef struct {
    float inv[3];
    int zero[3];
} ray_dir_pre;

static int ray_axis_update(float o, float lo, float hi, float inv, int zero, float *t_near, float *t_far)
{
    if (zero)
        return o >= lo && o <= hi;
    float t0 = (lo - o) * inv;
    float t1 = (hi - o) * inv;
    *t_near = fmax_std(*t_near, fmin_std(t0, t1));
    *t_far = fmin_std(*t_far, fmax_std(t0, t1));
    return *t_near <= *t_far;
}

/* gwn_ray_aabb_intersect_interval_impl (:100-140) */
static int ray_box_canonical(const float *o, const ray_dir_pre *d, const float *b, float t_min, float t_max,
                             float *t_near_out)
{
    if (!(t_max >= t_min))
        return 0;
    float t_near = t_min, t_far = t_max;
    for (int k = 0; k < 3; ++k)
        if (!ray_axis_update(o[k], b[k], b[3 + k], d->inv[k], d->zero[k], &t_near, &t_far))
            return 0;
    if (t_far < t_near)
        return 0;
    *t_near_out = t_near;
    return 1;
}

/* gwn_stable_triangle_normal_impl (:143-172) */
static void stable_normal(const float *a, const float *b, const float *c, float *n)
{
    float abx = a[2] * b[1], aby = a[0] * b[2], abz = a[1] * b[0];
    float bcx = b[2] * c[1], bcy = b[0] * c[2], bcz = b[1] * c[0];
    float cab[3] = {a[1] * b[2] - abx, a[2] * b[0] - aby, a[0] * b[1] - abz};
    float cbc[3] = {b[1] * c[2] - bcx, b[2] * c[0] - bcy, b[0] * c[1] - bcz};
    n[0] = cbc[0]; n[1] = cbc[1]; n[2] = cbc[2];
    if (fabsf(abx) < fabsf(bcx)) n[0] = cab[0];
    if (fabsf(aby) < fabsf(bcy)) n[1] = cab[1];
    if (fabsf(abz) < fabsf(bcz)) n[2] = cab[2];
}

static void cross3f(const float *a, const float *b, float *r)
{
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

/* gwn_ray_triangle_intersect_impl (:175-239): oriented Pluecker edge functions. */
static int ray_triangle(const float *o, const float *dir, const float *v0, const float *v1, const float *v2,
                        float t_min, float t_max, float *t_out, float *u_out, float *v_out, float *n_out)
{
    float p0[3], p1[3], p2[3], e0[3], e1[3], e2[3], s[3], c[3];
    for (int k = 0; k < 3; ++k) {
        p0[k] = v0[k] - o[k];
        p1[k] = v1[k] - o[k];
        p2[k] = v2[k] - o[k];
    }
    for (int k = 0; k < 3; ++k) {
        e0[k] = p2[k] - p0[k];
        e1[k] = p0[k] - p1[k];
        e2[k] = p1[k] - p2[k];
    }
    for (int k = 0; k < 3; ++k) s[k] = p2[k] + p0[k];
    cross3f(e0, s, c);
    float u = dot3f(c, dir);
    for (int k = 0; k < 3; ++k) s[k] = p0[k] + p1[k];
    cross3f(e1, s, c);
    float v = dot3f(c, dir);
    for (int k = 0; k < 3; ++k) s[k] = p1[k] + p2[k];
    cross3f(e2, s, c);
    float w = dot3f(c, dir);
    float uvw = u + v + w;
    float eps = FLT_EPSILON * fabsf(uvw);
    float mn = fmin_std(u, fmin_std(v, w));
    float mx = fmax_std(u, fmax_std(v, w));
    if (!(mn >= -eps || mx <= eps))
        return 0;
    float ng[3];
    stable_normal(e0, e1, e2, ng);
    float den = 2.0f * dot3f(ng, dir);
    if (den == 0.0f)
        return 0;
    float t_num = 2.0f * dot3f(p0, ng);
    float t = t_num / den;
    if (!(t_min <= t && t <= t_max))
        return 0;
    if (fabsf(uvw) >= FLT_MIN) {
        float inv = 1.0f / uvw;
        *u_out = u * inv;
        *v_out = v * inv;
    } else {
        if (uvw == 0.0f)
            return 0;
        *u_out = u / uvw;
        *v_out = v / uvw;
    }
    *t_out = t;
    n_out[0] = ng[0]; n_out[1] = ng[1]; n_out[2] = ng[2];
    return 1;
}

typedef struct {
    float t, u, v, n[3];
    uint32_t prim;
    int found;
} ray_best;

typedef struct {
    const gwn_oracle_node4_f32 *nodes;
    uint64_t node_count;
    gwn_oracle_child_f32 root;
    const float *tri;
    const uint32_t *prim;
    uint64_t nt;
    uint32_t packed_stack_bound;
    const float *ox, *oy, *oz, *dx, *dy, *dz;
    float t_min, t_max;
    int cap;
    int mode; /* -1: the batch's dispatch, 0: canonical, 1: packed */
    gwn_oracle_ray_hit *out;
} ray_ctx;

static void ray_visit_leaf(const ray_ctx *c, uint64_t ref, const float *o, const float *dir, float t_min,
                           ray_best *b)
{
    uint64_t off = ref & GWN_ORACLE_OFFSET_MASK;
    uint32_t n = (uint32_t)(ref >> 48);
    for (uint32_t k = 0; k < n; ++k) {
        if (off + k >= c->nt)
            continue;
        float a[3], bb[3], cc[3], t, u, v, ng[3];
        record_vertices_f32(c->tri + 9 * (off + k), a, bb, cc);
        if (!ray_triangle(o, dir, a, bb, cc, t_min, b->t, &t, &u, &v, ng))
            continue;
        if (b->found && !(t < b->t))
            continue;
        b->t = t; b->u = u; b->v = v;
        b->n[0] = ng[0]; b->n[1] = ng[1]; b->n[2] = ng[2];
        b->prim = c->prim[off + k];
        b->found = 1;
    }
}

#define CSWAPT(i, j)                                                                               \
    if (et[i] > et[j]) {                                                                           \
        float tt = et[i]; et[i] = et[j]; et[j] = tt;                                               \
        uint64_t tr = er[i]; er[i] = er[j]; er[j] = tr;                                            \
    }

/* Packed traversal (:242-358); returns 1 on stack overflow. */
static int ray_packed(const ray_ctx *c, const float *o, const float *dir, const ray_dir_pre *d, int parallel,
                      float t_min, ray_best *b)
{
    const uint32_t leaf_bit = 1u << 31;
    uint32_t stack[1024];
    int sp = 0;
    uint32_t payload = (uint32_t)(c->root.reference & GWN_ORACLE_OFFSET_MASK);
    for (;;) {
        if (payload & leaf_bit) {
            int slot = (int)(payload & 3u);
            uint32_t parent = (payload & ~leaf_bit) >> 2;
            ray_visit_leaf(c, c->nodes[parent].children[slot].reference, o, dir, t_min, b);
        } else if (payload < c->node_count) {
            const gwn_oracle_node4_f32 *nd = &c->nodes[payload];
            float et[4];
            uint64_t er[4];
            for (int s = 0; s < 4; ++s) {
                et[s] = INFINITY;
                er[s] = 0;
                const gwn_oracle_child_f32 *ch = &nd->children[s];
                uint64_t ref = ch->reference;
                if (!(ref & GWN_ORACLE_VALID_BIT))
                    continue;
                int leaf = (ref >> 48) != 0;
                if (!leaf && (ref & GWN_ORACLE_OFFSET_MASK) >= c->node_count)
                    continue;
                const float *bx = ch->bounds;
                float t_near, t_far;
                int hit;
                if (!parallel) {
                    float tx0 = (bx[0] - o[0]) * d->inv[0], tx1 = (bx[3] - o[0]) * d->inv[0];
                    float ty0 = (bx[1] - o[1]) * d->inv[1], ty1 = (bx[4] - o[1]) * d->inv[1];
                    float tz0 = (bx[2] - o[2]) * d->inv[2], tz1 = (bx[5] - o[2]) * d->inv[2];
                    t_near = fmax_std(t_min, fmax_std(fmin_std(tx0, tx1),
                                                      fmax_std(fmin_std(ty0, ty1), fmin_std(tz0, tz1))));
                    t_far = fmin_std(b->t, fmin_std(fmax_std(tx0, tx1),
                                                    fmin_std(fmax_std(ty0, ty1), fmax_std(tz0, tz1))));
                    hit = t_near <= t_far;
                } else {
                    t_near = t_min;
                    t_far = b->t;
                    hit = 1;
                    for (int k = 0; k < 3 && hit; ++k)
                        hit = ray_axis_update(o[k], bx[k], bx[3 + k], d->inv[k], d->zero[k], &t_near, &t_far);
                }
                if (!hit)
                    continue;
                et[s] = t_near;
                er[s] = leaf ? (leaf_bit | (payload << 2) | (uint32_t)s) : (uint32_t)(ref & GWN_ORACLE_OFFSET_MASK);
            }
            CSWAPT(0, 1) CSWAPT(2, 3) CSWAPT(0, 2) CSWAPT(1, 3) CSWAPT(1, 2)
            for (int s = 3; s > 0; --s) {
                if (et[s] == INFINITY)
                    continue;
                if (sp >= c->cap || sp >= 1024)
                    return 1;
                stack[sp++] = (uint32_t)er[s];
            }
            if (et[0] != INFINITY) {
                payload = (uint32_t)er[0];
                continue;
            }
        }
        if (sp == 0)
            return 0;
        payload = stack[--sp];
    }
}

/* Canonical traversal (:485-560); returns 1 on stack overflow. */
static int ray_canonical(const ray_ctx *c, const float *o, const float *dir, const ray_dir_pre *d, float t_min,
                         ray_best *b)
{
    uint64_t stack[1024];
    int sp = 0;
    uint64_t reference = c->root.reference;
    for (;;) {
        uint64_t next = 0;
        int valid = (reference & GWN_ORACLE_VALID_BIT) != 0;
        uint32_t count = (uint32_t)(reference >> 48);
        uint64_t off = reference & GWN_ORACLE_OFFSET_MASK;
        if (valid && count != 0) {
            ray_visit_leaf(c, reference, o, dir, t_min, b);
        } else if (valid && count == 0 && off < c->node_count) {
            const gwn_oracle_node4_f32 *nd = &c->nodes[off];
            float et[4];
            uint64_t er[4];
            int m = 0;
            for (int s = 0; s < 4; ++s) {
                const gwn_oracle_child_f32 *ch = &nd->children[s];
                if (!(ch->reference & GWN_ORACLE_VALID_BIT))
                    continue;
                float t_near;
                if (!ray_box_canonical(o, d, ch->bounds, t_min, b->t, &t_near))
                    continue;
                et[m] = t_near;
                er[m] = ch->reference;
                ++m;
            }
            for (int s = m; s < 4; ++s) {
                et[s] = INFINITY;
                er[s] = 0;
            }
            CSWAPT(0, 1) CSWAPT(2, 3) CSWAPT(0, 2) CSWAPT(1, 3) CSWAPT(1, 2)
            for (int s = 0; s < m; ++s)
                if ((er[s] & GWN_ORACLE_VALID_BIT) && (er[s] >> 48) != 0 && et[s] <= b->t)
                    ray_visit_leaf(c, er[s], o, dir, t_min, b);
            for (int s = m - 1; s >= 0; --s) {
                if (!((er[s] & GWN_ORACLE_VALID_BIT) && (er[s] >> 48) == 0) || et[s] > b->t)
                    continue;
                if (next != 0) {
                    if (sp >= c->cap || sp >= 1024)
                        return 1;
                    stack[sp++] = next;
                }
                next = er[s];
            }
        }
        if (next != 0) {
            reference = next;
            continue;
        }
        if (sp == 0)
            return 0;
        reference = stack[--sp];
    }
}
#undef CSWAPT

static void ray_one(const ray_ctx *c, uint64_t i, gwn_oracle_ray_hit *r)
{
    float o[3] = {c->ox[i], c->oy[i], c->oz[i]};
    float dir[3] = {c->dx[i], c->dy[i], c->dz[i]};
    memset(r, 0, sizeof(*r));
    r->t = -1.0f;
    r->primitive_id = 0xffffffffu;
    r->status = 0;
    float len2 = dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2];
    if (!(len2 > 0.0f))
        return;
    ray_dir_pre d;
    for (int k = 0; k < 3; ++k) {
        d.zero[k] = dir[k] == 0.0f;
        d.inv[k] = d.zero[k] ? 0.0f : 1.0f / dir[k];
    }
    ray_best b = {c->t_max, 0.0f, 0.0f, {0.0f, 0.0f, 0.0f}, 0xffffffffu, 0};
    int packed = (c->root.reference & GWN_ORACLE_VALID_BIT) && (c->root.reference >> 48) == 0 &&
                 c->node_count < (1ull << 29) && (int)c->packed_stack_bound <= c->cap;
    if (c->mode >= 0)
        packed = c->mode == 1 && (c->root.reference >> 48) == 0;
    int overflow = packed ? ray_packed(c, o, dir, &d, d.zero[0] || d.zero[1] || d.zero[2], c->t_min, &b)
                          : ray_canonical(c, o, dir, &d, c->t_min, &b);
    if (b.found) {
        r->t = b.t; r->u = b.u; r->v = b.v;
        r->nx = b.n[0]; r->ny = b.n[1]; r->nz = b.n[2];
        r->primitive_id = b.prim;
        r->status = 1;
    }
    if (overflow)
        r->status = 2;
}

static void ray_range(void *vctx, uint64_t begin, uint64_t end)
{
    ray_ctx *c = (ray_ctx *)vctx;
    for (uint64_t i = begin; i < end; ++i)
        ray_one(c, i, &c->out[i]);
}

void gwn_oracle_ray_first_hit_f32(
    const gwn_oracle_node4_f32 *nodes, uint64_t node_count, gwn_oracle_child_f32 root,
    const float *triangles, const uint32_t *primitive_indices, uint64_t triangle_count,
    uint32_t packed_stack_bound, const float *ox, const float *oy, const float *oz, const float *dx,
    const float *dy, const float *dz, uint64_t ray_count, float t_min, float t_max, int stack_capacity,
    int traversal, gwn_oracle_ray_hit *out, int threads)
{
    ray_ctx ctx = {nodes, node_count, root, triangles, primitive_indices, triangle_count, packed_stack_bound,
                   ox, oy, oz, dx, dy, dz, t_min, t_max, stack_capacity, traversal, out};
    parallel_ranges(ray_range, &ctx, ray_count, threads);
}

/* ---------------------------------------------------------------------------
 * Boundary chain, include/gwn/detail/gwn_boundary_impl.cuh:21-354: every oriented triangle edge
 * becomes an undirected (lo, hi) key with orientation +1 (lo -> hi), -1 (hi -> lo) or 0 (equal
 * endpoints); keys are ordered by (lo, hi); each key group with a nonzero orientation sum emits
 * one directed row (lo -> hi for a positive sum, hi -> lo for a negative one) with multiplicity
 * |sum|, in key order.
 * ------------------------------------------------------------------------- */

typedef struct {
    uint64_t key; /* lo << 32 | hi */
    int orientation;
} boundary_edge;

static int cmp_boundary_edge(const void *pa, const void *pb)
{
    const boundary_edge *a = (const boundary_edge *)pa, *b = (const boundary_edge *)pb;
    return (a->key > b->key) - (a->key < b->key);
}

int gwn_oracle_boundary_chain(
    const uint32_t *i0, const uint32_t *i1, const uint32_t *i2, uint64_t triangle_count,
    uint32_t *start, uint32_t *end, uint64_t *multiplicity, uint64_t *row_count)
{
    uint64_t m = 3 * triangle_count;
    boundary_edge *e = (boundary_edge *)malloc(sizeof(boundary_edge) * (m ? m : 1));
    if (!e)
        return GWN_ORACLE_INTERNAL_ERROR;
    for (uint64_t t = 0; t < triangle_count; ++t) {
        uint32_t v[3] = {i0[t], i1[t], i2[t]};
        for (int k = 0; k < 3; ++k) {
            uint32_t a = v[k], b = v[(k + 1) % 3];
            uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
            e[3 * t + k].key = ((uint64_t)lo << 32) | hi;
            e[3 * t + k].orientation = a < b ? 1 : (b < a ? -1 : 0);
        }
    }
    qsort(e, m, sizeof(boundary_edge), cmp_boundary_edge);
    uint64_t rows = 0;
    for (uint64_t i = 0; i < m;) {
        uint64_t j = i;
        long long sum = 0;
        while (j < m && e[j].key == e[i].key)
            sum += e[j++].orientation;
        if (sum != 0) {
            uint32_t lo = (uint32_t)(e[i].key >> 32), hi = (uint32_t)e[i].key;
            if (start) {
                start[rows] = sum > 0 ? lo : hi;
                end[rows] = sum > 0 ? hi : lo;
                multiplicity[rows] = (uint64_t)(sum > 0 ? sum : -sum);
            }
            ++rows;
        }
        i = j;
    }
    *row_count = rows;
    free(e);
    return GWN_ORACLE_OK;
}

/* ---------------------------------------------------------------------------
 * Antipodal winding number and gradient, include/gwn/detail/
 * gwn_query_winding_antipodal_impl.cuh:55-552: signed crossings of a positive coordinate ray
 * with the triangles, plus the signed spherical areas of the boundary edges seen from the
 * query; a singular axis (projected edge through the query, triangle through the query) restarts
 * both terms on the next axis in the order Z, X, Y; NaN when all three are singular.
 * near_fp64 != 0 (precision split) evaluates each boundary area / area gradient in FP64 on the
 * exact vertices and sums in FP64; classification and crossings stay in FP32.
 * ------------------------------------------------------------------------- */

enum { AXIS_X = 0, AXIS_Y = 1, AXIS_Z = 2 };

static void project2(const float *v, int axis, float *p)
{
    if (axis == AXIS_X) { p[0] = v[1]; p[1] = v[2]; }
    else if (axis == AXIS_Y) { p[0] = v[2]; p[1] = v[0]; }
    else { p[0] = v[0]; p[1] = v[1]; }
}

static int sign_of(float v) { return v > 0.0f ? 1 : (v < 0.0f ? -1 : 0); }

/* (:64-110) returns the perturbed sign; *singular when the projected segment holds the query. */
static int classify_edge(const float *s, const float *e, int axis, int *singular, float *det)
{
    float a[2], b[2];
    project2(s, axis, a);
    project2(e, axis, b);
    float d = a[0] * b[1] - a[1] * b[0];
    int sg = sign_of(d);
    if (sg == 0) {
        sg = sign_of(a[1] - b[1]);
        if (sg == 0)
            sg = sign_of(b[0] - a[0]);
    }
    *singular = d == 0.0f && (a[0] * b[0] + a[1] * b[1] <= 0.0f);
    *det = d;
    return sg;
}

/* (:122-176) returns 1 when singular; *crossing receives +-1 or stays 0. */
static int triangle_crossing(const float *o, int axis, const float *a, const float *b, const float *c, float *crossing)
{
    float pa[3], pb[3], pc[3], u[3], v[3], n[3];
    for (int k = 0; k < 3; ++k) {
        pa[k] = a[k] - o[k];
        pb[k] = b[k] - o[k];
        pc[k] = c[k] - o[k];
    }
    for (int k = 0; k < 3; ++k) {
        u[k] = pb[k] - pa[k];
        v[k] = pc[k] - pa[k];
    }
    cross3f(u, v, n);
    float na = n[axis];
    if (na == 0.0f)
        return 0;
    int s1, s2, s3;
    float d;
    int e1 = classify_edge(pa, pb, axis, &s1, &d);
    int e2 = classify_edge(pb, pc, axis, &s2, &d);
    int e3 = classify_edge(pc, pa, axis, &s3, &d);
    if (s1 || s2 || s3)
        return 1;
    if (e1 == 0 || e2 == 0 || e3 == 0)
        return 1;
    if (!((e1 >= 0 && e2 >= 0 && e3 >= 0) || (e1 <= 0 && e2 <= 0 && e3 <= 0)))
        return 0;
    float t = dot3f(n, pa) / na;
    if (t == 0.0f)
        return 1;
    if (!(t > 0.0f))
        return 0;
    *crossing = na > 0.0f ? 1.0f : -1.0f;
    return 0;
}

typedef struct {
    const gwn_oracle_node4_f32 *nodes;
    uint64_t node_count;
    gwn_oracle_child_f32 root;
    const float *tri;
    uint64_t nt;
    const float *vx, *vy, *vz;
    const uint32_t *bs, *be;
    const uint64_t *bm;
    uint64_t n_edges;
    const float *qx, *qy, *qz;
    int cap;
    int near_fp64;
    float *out;
    float *gx, *gy, *gz;
} antipodal_ctx;

static int positive_axis_hits_box(const float *o, int axis, const float *b)
{
    if (axis == AXIS_X)
        return o[1] >= b[1] && o[1] <= b[4] && o[2] >= b[2] && o[2] <= b[5] && b[3] >= o[0];
    if (axis == AXIS_Y)
        return o[2] >= b[2] && o[2] <= b[5] && o[0] >= b[0] && o[0] <= b[3] && b[4] >= o[1];
    return o[0] >= b[0] && o[0] <= b[3] && o[1] >= b[1] && o[1] <= b[4] && b[5] >= o[2];
}

/* Returns 1 when singular; *sum receives the signed crossing count (NaN on stack overflow). */
static int leaf_crossings(const antipodal_ctx *c, uint64_t ref, const float *o, int axis, float *sum)
{
    uint64_t off = ref & GWN_ORACLE_OFFSET_MASK;
    uint32_t n = (uint32_t)(ref >> 48);
    for (uint32_t k = 0; k < n; ++k) {
        if (off + k >= c->nt)
            continue;
        float a[3], b[3], cc[3], cr = 0.0f;
        record_vertices_f32(c->tri + 9 * (off + k), a, b, cc);
        if (triangle_crossing(o, axis, a, b, cc, &cr))
            return 1;
        *sum += cr;
    }
    return 0;
}

static int crossing_count(const antipodal_ctx *c, const float *o, int axis, float *sum)
{
    *sum = 0.0f;
    if ((c->root.reference & GWN_ORACLE_VALID_BIT) && (c->root.reference >> 48) != 0)
        return leaf_crossings(c, c->root.reference, o, axis, sum);
    uint32_t stack[1024];
    int sp = 0;
    uint32_t node = (uint32_t)(c->root.reference & GWN_ORACLE_OFFSET_MASK);
    for (;;) {
        uint32_t next = GWN_ORACLE_INVALID_INDEX;
        if (node < c->node_count) {
            const gwn_oracle_node4_f32 *nd = &c->nodes[node];
            for (int s = 0; s < 4; ++s) {
                const gwn_oracle_child_f32 *ch = &nd->children[s];
                if (!(ch->reference & GWN_ORACLE_VALID_BIT))
                    continue;
                if (!positive_axis_hits_box(o, axis, ch->bounds))
                    continue;
                if ((ch->reference >> 48) != 0) {
                    if (leaf_crossings(c, ch->reference, o, axis, sum))
                        return 1;
                    continue;
                }
                if (next != GWN_ORACLE_INVALID_INDEX) {
                    if (sp >= c->cap || sp >= 1024) {
                        *sum = NAN;
                        return 0;
                    }
                    stack[sp++] = next;
                }
                next = (uint32_t)(ch->reference & GWN_ORACLE_OFFSET_MASK);
            }
        }
        if (next != GWN_ORACLE_INVALID_INDEX) {
            node = next;
            continue;
        }
        if (sp == 0)
            return 0;
        node = stack[--sp];
    }
}

static void minus_ray(int axis, float *m)
{
    m[0] = -(axis == AXIS_X ? 1.0f : 0.0f);
    m[1] = -(axis == AXIS_Y ? 1.0f : 0.0f);
    m[2] = -(axis == AXIS_Z ? 1.0f : 0.0f);
}

/* (:178-204) */
static float spherical_area_f32(const float *m, const float *s, const float *e, float det)
{
    float sn = sqrtf(dot3f(s, s)), en = sqrtf(dot3f(e, e));
    if (!(sn > 0.0f) || !(en > 0.0f))
        return NAN;
    float num = -det;
    if (det == 0.0f) {
        float cr[3];
        cross3f(s, e, cr);
        num = dot3f(m, cr);
    }
    float den = sn * en + dot3f(m, s) * en + dot3f(s, e) + dot3f(e, m) * sn;
    return 2.0f * (float)atan2((double)num, (double)den);
}

static double dot3d(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void cross3d(const double *a, const double *b, double *r)
{
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

/* FP64 area of the same term (precision split); the branch sign comes from the FP32 class. */
static double spherical_area_f64(const double *m, const double *s, const double *e)
{
    double sn = sqrt(dot3d(s, s)), en = sqrt(dot3d(e, e));
    if (!(sn > 0.0) || !(en > 0.0))
        return NAN;
    double cr[3];
    cross3d(s, e, cr);
    double num = dot3d(m, cr);
    double den = sn * en + dot3d(m, s) * en + dot3d(s, e) + dot3d(e, m) * sn;
    return 2.0 * atan2(num, den);
}

/* (:210-255) returns 1 when singular. */
static int boundary_term(const antipodal_ctx *c, const float *q, int axis, float *value, double *value_d)
{
    const float pi = (float)3.141592653589793238462643383279502884L;
    const double pi_d = 3.141592653589793238462643383279502884;
    float m[3];
    minus_ray(axis, m);
    double md[3] = {m[0], m[1], m[2]};
    float sum = 0.0f, comp = 0.0f;
    double sum_d = 0.0;
    for (uint64_t i = 0; i < c->n_edges; ++i) {
        uint32_t a = c->bs[i], b = c->be[i];
        float s[3] = {c->vx[a] - q[0], c->vy[a] - q[1], c->vz[a] - q[2]};
        float e[3] = {c->vx[b] - q[0], c->vy[b] - q[1], c->vz[b] - q[2]};
        int singular;
        float det;
        int sg = classify_edge(s, e, axis, &singular, &det);
        if (singular)
            return 1;
        if (c->near_fp64) {
            double sd[3] = {(double)c->vx[a] - q[0], (double)c->vy[a] - q[1], (double)c->vz[a] - q[2]};
            double ed[3] = {(double)c->vx[b] - q[0], (double)c->vy[b] - q[1], (double)c->vz[b] - q[2]};
            double area = spherical_area_f64(md, sd, ed);
            if (axis != AXIS_Z && sg > 0 && area < 0.0)
                area += 2.0 * pi_d;
            sum_d += (double)c->bm[i] * area;
            continue;
        }
        float area = spherical_area_f32(m, s, e, det);
        if (axis != AXIS_Z && sg > 0 && area < 0.0f)
            area += 2.0f * pi;
        float term = (float)c->bm[i] * area;
        float ct = term - comp;
        float next = sum + ct;
        comp = (next - sum) - ct;
        sum = next;
    }
    *value = sum / (4.0f * pi);
    *value_d = sum_d / (4.0 * pi_d);
    return 0;
}

static void antipodal_one(const antipodal_ctx *c, uint64_t qi)
{
    float q[3] = {c->qx[qi], c->qy[qi], c->qz[qi]};
    const int order[3] = {AXIS_Z, AXIS_X, AXIS_Y};
    float result = NAN;
    for (int k = 0; k < 3; ++k) {
        float crossings;
        if (crossing_count(c, q, order[k], &crossings))
            continue;
        float b;
        double bd;
        if (boundary_term(c, q, order[k], &b, &bd))
            continue;
        result = c->near_fp64 ? (float)((double)crossings + bd) : crossings + b;
        break;
    }
    c->out[qi] = result;
}

/* (:257-334) */
static void area_gradient_f32(const float *m, const float *s, const float *e, float det, float *g)
{
    float sn = sqrtf(dot3f(s, s)), en = sqrtf(dot3f(e, e));
    if (sn == 0.0f || en == 0.0f) {
        g[0] = g[1] = g[2] = NAN;
        return;
    }
    float is = 1.0f / sn, ie = 1.0f / en;
    float su[3] = {s[0] * is, s[1] * is, s[2] * is};
    float eu[3] = {e[0] * ie, e[1] * ie, e[2] * ie};
    float num = -det;
    if (det == 0.0f) {
        float cr[3];
        cross3f(s, e, cr);
        num = dot3f(m, cr);
    }
    float den = sn * en + dot3f(m, s) * en + dot3f(m, e) * sn + dot3f(s, e);
    float d2 = num * num + den * den;
    if (d2 == 0.0f) {
        g[0] = g[1] = g[2] = NAN;
        return;
    }
    float cn = 2.0f * den / d2;
    float cd = -2.0f * num / d2;
    float c1[3], c2[3], gn[3], t1[3], t2[3], gd[3];
    cross3f(e, m, c1);
    cross3f(m, s, c2);
    for (int k = 0; k < 3; ++k)
        gn[k] = -(c1[k] + c2[k]);
    float f1 = 1.0f + dot3f(m, eu);
    float f2 = 1.0f + dot3f(m, su);
    for (int k = 0; k < 3; ++k) {
        t1[k] = en * (m[k] + eu[k] + f1 * su[k]);
        t2[k] = sn * (m[k] + su[k] + f2 * eu[k]);
        gd[k] = -(t1[k] + t2[k]);
    }
    for (int k = 0; k < 3; ++k)
        g[k] = cn * gn[k] + cd * gd[k];
}

static void area_gradient_f64(const double *m, const double *s, const double *e, double *g)
{
    double sn = sqrt(dot3d(s, s)), en = sqrt(dot3d(e, e));
    if (sn == 0.0 || en == 0.0) {
        g[0] = g[1] = g[2] = NAN;
        return;
    }
    double su[3] = {s[0] / sn, s[1] / sn, s[2] / sn};
    double eu[3] = {e[0] / en, e[1] / en, e[2] / en};
    double cr[3];
    cross3d(s, e, cr);
    double num = dot3d(m, cr);
    double den = sn * en + dot3d(m, s) * en + dot3d(m, e) * sn + dot3d(s, e);
    double d2 = num * num + den * den;
    if (d2 == 0.0) {
        g[0] = g[1] = g[2] = NAN;
        return;
    }
    double cn = 2.0 * den / d2, cd = -2.0 * num / d2;
    double c1[3], c2[3];
    cross3d(e, m, c1);
    cross3d(m, s, c2);
    double f1 = 1.0 + dot3d(m, eu), f2 = 1.0 + dot3d(m, su);
    for (int k = 0; k < 3; ++k) {
        double gn = -(c1[k] + c2[k]);
        double gdk = -(en * (m[k] + eu[k] + f1 * su[k]) + sn * (m[k] + su[k] + f2 * eu[k]));
        g[k] = cn * gn + cd * gdk;
    }
}

/* (:337-376) returns 1 when singular. */
static int boundary_gradient(const antipodal_ctx *c, const float *q, int axis, float *out3)
{
    const float pi = (float)3.141592653589793238462643383279502884L;
    const double pi_d = 3.141592653589793238462643383279502884;
    float m[3];
    minus_ray(axis, m);
    double md[3] = {m[0], m[1], m[2]};
    float sum[3] = {0.0f, 0.0f, 0.0f}, comp[3] = {0.0f, 0.0f, 0.0f};
    double sum_d[3] = {0.0, 0.0, 0.0};
    for (uint64_t i = 0; i < c->n_edges; ++i) {
        uint32_t a = c->bs[i], b = c->be[i];
        float s[3] = {c->vx[a] - q[0], c->vy[a] - q[1], c->vz[a] - q[2]};
        float e[3] = {c->vx[b] - q[0], c->vy[b] - q[1], c->vz[b] - q[2]};
        int singular;
        float det;
        classify_edge(s, e, axis, &singular, &det);
        if (singular)
            return 1;
        if (c->near_fp64) {
            double sd[3] = {(double)c->vx[a] - q[0], (double)c->vy[a] - q[1], (double)c->vz[a] - q[2]};
            double ed[3] = {(double)c->vx[b] - q[0], (double)c->vy[b] - q[1], (double)c->vz[b] - q[2]};
            double g[3];
            area_gradient_f64(md, sd, ed, g);
            for (int k = 0; k < 3; ++k)
                sum_d[k] += (double)c->bm[i] * g[k];
            continue;
        }
        float g[3];
        area_gradient_f32(m, s, e, det, g);
        float mult = (float)c->bm[i];
        for (int k = 0; k < 3; ++k) {
            float term = mult * g[k];
            float ct = term - comp[k];
            float next = sum[k] + ct;
            comp[k] = (next - sum[k]) - ct;
            sum[k] = next;
        }
    }
    for (int k = 0; k < 3; ++k)
        out3[k] = c->near_fp64 ? (float)(sum_d[k] / (4.0 * pi_d)) : sum[k] / (4.0f * pi);
    return 0;
}

static void antipodal_gradient_one(const antipodal_ctx *c, uint64_t qi)
{
    float q[3] = {c->qx[qi], c->qy[qi], c->qz[qi]};
    const int order[3] = {AXIS_Z, AXIS_X, AXIS_Y};
    float g[3] = {NAN, NAN, NAN};
    for (int k = 0; k < 3; ++k)
        if (!boundary_gradient(c, q, order[k], g))
            break;
        else
            g[0] = g[1] = g[2] = NAN;
    c->gx[qi] = g[0];
    c->gy[qi] = g[1];
    c->gz[qi] = g[2];
}

static void antipodal_range(void *vctx, uint64_t begin, uint64_t end)
{
    antipodal_ctx *c = (antipodal_ctx *)vctx;
    for (uint64_t q = begin; q < end; ++q) {
        if (c->out)
            antipodal_one(c, q);
        if (c->gx)
            antipodal_gradient_one(c, q);
    }
}

void gwn_oracle_winding_antipodal_f32(
    const gwn_oracle_node4_f32 *nodes, uint64_t node_count, gwn_oracle_child_f32 root,
    const float *triangles, uint64_t triangle_count, const float *vx, const float *vy, const float *vz,
    const uint32_t *boundary_start, const uint32_t *boundary_end, const uint64_t *boundary_multiplicity,
    uint64_t boundary_edges, const float *qx, const float *qy, const float *qz, uint64_t query_count,
    int stack_capacity, int near_fp64, float *out, float *gx, float *gy, float *gz, int threads)
{
    antipodal_ctx ctx = {nodes, node_count, root, triangles, triangle_count, vx, vy, vz, boundary_start,
                         boundary_end, boundary_multiplicity, boundary_edges, qx, qy, qz, stack_capacity,
                         near_fp64, out, gx, gy, gz};
    parallel_ranges(antipodal_range, &ctx, query_count, threads);
}

/* ---------------------------------------------------------------------------
 * Adjoint of winding-number and distance losses (no reference counterpart; pinned by FP64 central
 * differences of gwn_oracle_winding_taylor_frozen_f64 and brute-force distances).
 *
 * L = sum_q u_q w_q for the Taylor winding (orders 0, 1) and L = sum_p u_p d_p for the unsigned
 * distance.  The Taylor replay takes the forward's accept/descend decisions (FP32 fields, strict
 * |q - P|^2 > maxPDist2 * beta^2 and a finite term) and differentiates the chosen terms in FP64:
 *   far term of child slot s:  w_s(q - P_s, N_s, T_s), T_s the first-order fields;
 *   slot fields from the slot's additive payload (area A, A P, N, N_i c_j):
 *     P = AP / A,  T_ii = (N c)_ii - P_i N_i,  S_ij = (N c)_ij - P_j N_i + (N c)_ji - P_i N_j;
 *   payload of triangle (a, b, c): N = (b - a) x (c - a) / 2, A = |N|, c = (a + b + c) / 3;
 *   near triangles: W = 2 atan2(a.(b x c), |a||b||c| + (a.b)|c| + (a.c)|b| + (b.c)|a|) with
 *     a, b, c relative to q.
 * Every slot's payload adjoint reaches each triangle below it (the payloads are sums), so a
 * root-to-leaf walk adds the slot adjoints along each leaf's path.
 * ------------------------------------------------------------------------- */

typedef struct {
    double area, ap[3], n[3], n1[9];
} adj_payload;

static void adj_payload_add(adj_payload *d, const adj_payload *s)
{
    d->area += s->area;
    for (int k = 0; k < 3; ++k) {
        d->ap[k] += s->ap[k];
        d->n[k] += s->n[k];
    }
    for (int k = 0; k < 9; ++k)
        d->n1[k] += s->n1[k];
}

static void adj_tri_payload(const double *a, const double *b, const double *c, adj_payload *p)
{
    double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double n[3];
    cross3d(e1, e2, n);
    for (int k = 0; k < 3; ++k)
        n[k] *= 0.5;
    double cen[3] = {(a[0] + b[0] + c[0]) / 3.0, (a[1] + b[1] + c[1]) / 3.0, (a[2] + b[2] + c[2]) / 3.0};
    p->area = sqrt(dot3d(n, n));
    for (int k = 0; k < 3; ++k) {
        p->ap[k] = cen[k] * p->area;
        p->n[k] = n[k];
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            p->n1[3 * i + j] = n[i] * cen[j];
}

/* Vertices of leaf record k in FP64 from the vertex arrays (primitive order + indices). */
typedef struct {
    const double *vx, *vy, *vz;
    const uint32_t *prim, *i0, *i1, *i2;
} adj_mesh;

static void adj_record_vertices(const adj_mesh *m, uint64_t k, double *a, double *b, double *c, uint32_t *ids)
{
    uint32_t p = m->prim[k];
    ids[0] = m->i0[p];
    ids[1] = m->i1[p];
    ids[2] = m->i2[p];
    double *v[3] = {a, b, c};
    for (int t = 0; t < 3; ++t) {
        v[t][0] = m->vx[ids[t]];
        v[t][1] = m->vy[ids[t]];
        v[t][2] = m->vz[ids[t]];
    }
}

/* FP64 payload of every child slot (node * 4 + slot), children before parents. */
static void adj_slot_payloads(const gwn_oracle_node4_f32 *nodes, uint64_t node_count, const adj_mesh *m,
                              adj_payload *slot)
{
    for (uint64_t n = node_count; n-- > 0;) {
        for (int s = 0; s < 4; ++s) {
            adj_payload *out = &slot[4 * n + s];
            memset(out, 0, sizeof(*out));
            uint64_t ref = nodes[n].children[s].reference;
            if (!(ref & GWN_ORACLE_VALID_BIT))
                continue;
            uint64_t off = ref & GWN_ORACLE_OFFSET_MASK;
            uint32_t cnt = (uint32_t)(ref >> 48);
            if (cnt == 0) {
                for (int t = 0; t < 4; ++t)
                    adj_payload_add(out, &slot[4 * off + t]);
                continue;
            }
            for (uint32_t k = 0; k < cnt; ++k) {
                double a[3], b[3], c[3];
                uint32_t ids[3];
                adj_record_vertices(m, off + k, a, b, c, ids);
                adj_payload tp;
                adj_tri_payload(a, b, c, &tp);
                adj_payload_add(out, &tp);
            }
        }
    }
}

/* Slot fields P (3), N (3), first-order T (6) in FP64; zero-area slots keep the box centre. */
static void adj_slot_fields(const adj_payload *p, const float *bounds, double *P, double *N, double *T)
{
    for (int k = 0; k < 3; ++k) {
        P[k] = p->area > 0.0 ? p->ap[k] / p->area : ((double)bounds[k] + (double)bounds[3 + k]) * 0.5;
        N[k] = p->n[k];
    }
    const double *f = p->n1;
    T[0] = f[0] - P[0] * N[0];
    T[1] = f[4] - P[1] * N[1];
    T[2] = f[8] - P[2] * N[2];
    T[3] = (f[1] - P[1] * N[0]) + (f[3] - P[0] * N[1]);
    T[4] = (f[5] - P[2] * N[1]) + (f[7] - P[1] * N[2]);
    T[5] = (f[6] - P[0] * N[2]) + (f[2] - P[2] * N[0]);
}

/* Far term (orders 0, 1) at r = q - P and its derivatives: dw/dr (3), dw/dN (3), dw/dT (6). */
static double adj_far_term(const double *r, const double *N, const double *T, int order, double *dr, double *dN,
                           double *dT)
{
    double l2 = dot3d(r, r), m1 = 1.0 / sqrt(l2), m2 = m1 * m1, m3 = m2 * m1, m4 = m2 * m2;
    double qn[3] = {r[0] * m1, r[1] * m1, r[2] * m1};
    double rdotN = dot3d(qn, N);
    double w = -m2 * rdotN;
    for (int k = 0; k < 3; ++k) {
        dr[k] = m3 * (3.0 * qn[k] * rdotN - N[k]);
        dN[k] = -m2 * qn[k];
    }
    for (int k = 0; k < 6; ++k)
        dT[k] = 0.0;
    if (order >= 1) {
        double Tr = T[0] + T[1] + T[2];
        double S = qn[0] * qn[0] * T[0] + qn[1] * qn[1] * T[1] + qn[2] * qn[2] * T[2] + qn[0] * qn[1] * T[3] +
                   qn[1] * qn[2] * T[4] + qn[2] * qn[0] * T[5];
        w += m3 * (Tr - 3.0 * S);
        double v[3] = {2.0 * T[0] * qn[0] + T[3] * qn[1] + T[5] * qn[2],
                       T[3] * qn[0] + 2.0 * T[1] * qn[1] + T[4] * qn[2],
                       T[5] * qn[0] + T[4] * qn[1] + 2.0 * T[2] * qn[2]};
        double coeff = 15.0 * S - 3.0 * Tr;
        for (int k = 0; k < 3; ++k)
            dr[k] += m4 * (coeff * qn[k] - 3.0 * v[k]);
        for (int k = 0; k < 3; ++k)
            dT[k] = m3 * (1.0 - 3.0 * qn[k] * qn[k]);
        dT[3] = -3.0 * m3 * qn[0] * qn[1];
        dT[4] = -3.0 * m3 * qn[1] * qn[2];
        dT[5] = -3.0 * m3 * qn[2] * qn[0];
    }
    return w;
}

/* Solid angle of (a, b, c) seen from q (unnormalised atan2 form) and its vertex derivatives. */
static double adj_solid_angle(const double *va, const double *vb, const double *vc, const double *q, double *da,
                              double *db, double *dc)
{
    double a[3], b[3], c[3];
    for (int k = 0; k < 3; ++k) {
        a[k] = va[k] - q[k];
        b[k] = vb[k] - q[k];
        c[k] = vc[k] - q[k];
        da[k] = db[k] = dc[k] = 0.0;
    }
    double la = sqrt(dot3d(a, a)), lb = sqrt(dot3d(b, b)), lc = sqrt(dot3d(c, c));
    if (la == 0.0 || lb == 0.0 || lc == 0.0)
        return 0.0;
    double bc[3], ca[3], ab[3];
    cross3d(b, c, bc);
    cross3d(c, a, ca);
    cross3d(a, b, ab);
    double num = dot3d(a, bc);
    if (num == 0.0)
        return 0.0;
    double dab = dot3d(a, b), dac = dot3d(a, c), dbc = dot3d(b, c);
    double den = la * lb * lc + dab * lc + dac * lb + dbc * la;
    double s = 2.0 / (num * num + den * den);
    for (int k = 0; k < 3; ++k) {
        double ha = a[k] / la, hb = b[k] / lb, hc = c[k] / lc;
        double dDa = ha * lb * lc + b[k] * lc + c[k] * lb + dbc * ha;
        double dDb = hb * la * lc + a[k] * lc + c[k] * la + dac * hb;
        double dDc = hc * la * lb + a[k] * lb + b[k] * la + dab * hc;
        da[k] = s * (den * bc[k] - num * dDa);
        db[k] = s * (den * ca[k] - num * dDb);
        dc[k] = s * (den * ab[k] - num * dDc);
    }
    return 2.0 * atan2(num, den);
}

typedef struct {
    const gwn_oracle_node4_f32 *nodes;
    uint64_t node_count;
    gwn_oracle_child_f32 root;
    const float *tri;          /* FP32 records: the forward's decisions */
    uint64_t nt;
    const float *mom;          /* FP32 fields: the forward's decisions */
    int order, F;
    float beta;
    const adj_mesh *mesh;      /* FP64 values */
    const adj_payload *slot;
    const float *qx, *qy, *qz;
    const double *u;
    int threads;
    /* per-thread outputs */
    double *dslot;             /* 12 per slot: dP, dN, dT */
    double *dtri;              /* 9 per record */
    double *dq;                /* 3 per query (disjoint rows) */
    double value;              /* forward value sum_q u_q w_q (frozen path) */
} adj_ctx;

typedef struct {
    const adj_ctx *c;
    uint64_t begin, end;
    double *dslot, *dtri;
    double value;
} adj_job;

/* Replays one query's decisions; with grad != 0 accumulates derivatives, always returns w_q. */
static double adj_query(const adj_ctx *c, uint64_t qi, double *dslot, double *dtri, double *dqrow, int grad)
{
    const double inv4pi = 1.0 / (4.0 * 3.141592653589793238462643383279502884);
    float qf[3] = {c->qx[qi], c->qy[qi], c->qz[qi]};
    double q[3] = {qf[0], qf[1], qf[2]};
    double u = grad ? c->u[qi] * inv4pi : 0.0;
    double w = 0.0;
    double dq[3] = {0.0, 0.0, 0.0};
    float beta2 = c->beta * c->beta;
    uint32_t stack[1024];
    int sp = 0;
#define ADJ_LEAF(ref_)                                                                                 \
    do {                                                                                               \
        uint64_t off_ = (ref_) & GWN_ORACLE_OFFSET_MASK;                                              \
        uint32_t n_ = (uint32_t)((ref_) >> 48);                                                       \
        for (uint32_t k_ = 0; k_ < n_; ++k_) {                                                        \
            if (off_ + k_ >= c->nt)                                                                   \
                continue;                                                                             \
            double a_[3], b_[3], cc_[3], ga_[3], gb_[3], gc_[3];                                      \
            uint32_t ids_[3];                                                                         \
            adj_record_vertices(c->mesh, off_ + k_, a_, b_, cc_, ids_);                               \
            w += adj_solid_angle(a_, b_, cc_, q, ga_, gb_, gc_);                                      \
            if (grad) {                                                                               \
                double *t_ = dtri + 9 * (off_ + k_);                                                  \
                for (int d_ = 0; d_ < 3; ++d_) {                                                      \
                    t_[d_] += u * ga_[d_];                                                            \
                    t_[3 + d_] += u * gb_[d_];                                                        \
                    t_[6 + d_] += u * gc_[d_];                                                        \
                    dq[d_] -= u * (ga_[d_] + gb_[d_] + gc_[d_]);                                      \
                }                                                                                     \
            }                                                                                         \
        }                                                                                             \
    } while (0)
    if ((c->root.reference & GWN_ORACLE_VALID_BIT) && (c->root.reference >> 48) != 0) {
        ADJ_LEAF(c->root.reference);
    } else {
        uint32_t node = (uint32_t)(c->root.reference & GWN_ORACLE_OFFSET_MASK);
        for (;;) {
            if (node >= c->node_count)
                break;
            const gwn_oracle_node4_f32 *nd = &c->nodes[node];
            const float *t = c->mom + (size_t)node * 4 * c->F;
            uint32_t next = GWN_ORACLE_INVALID_INDEX;
            for (int s = 0; s < 4; ++s) {
                const gwn_oracle_child_f32 *ch = &nd->children[s];
                if (!(ch->reference & GWN_ORACLE_VALID_BIT))
                    continue;
#define TF(f) t[(f) * 4 + s]
                /* The forward's decision, in FP32 (gwn_query_winding_impl.cuh:330-364). */
                float qrx = qf[0] - TF(1), qry = qf[1] - TF(2), qrz = qf[2] - TF(3);
                float ql2 = qrx * qrx + qry * qry + qrz * qrz;
                int descend = !(ql2 > 0.0f) || ql2 <= TF(0) * beta2;
                if (!descend) {
                    float m1 = 1.0f / sqrtf(ql2), m2 = m1 * m1, m3 = m2 * m1;
                    float qn[3] = {qrx * m1, qry * m1, qrz * m1};
                    float fw = -m2 * (qn[0] * TF(4) + qn[1] * TF(5) + qn[2] * TF(6));
                    if (c->order >= 1) {
                        float Tr = TF(7) + TF(8) + TF(9);
                        float S = qn[0] * qn[0] * TF(7) + qn[1] * qn[1] * TF(8) + qn[2] * qn[2] * TF(9) +
                                  qn[0] * qn[1] * TF(10) + qn[0] * qn[2] * TF(12) + qn[1] * qn[2] * TF(11);
                        fw += m3 * (Tr - 3.0f * S);
                    }
                    descend = !isfinite(fw);
                }
#undef TF
                if (!descend) {
                    double P[3], N[3], T[6], r[3], dr[3], dN[3], dT[6];
                    adj_slot_fields(&c->slot[4 * node + s], ch->bounds, P, N, T);
                    for (int k = 0; k < 3; ++k)
                        r[k] = q[k] - P[k];
                    w += adj_far_term(r, N, T, c->order, dr, dN, dT);
                    if (grad) {
                        double *g = dslot + 12 * ((size_t)node * 4 + s);
                        for (int k = 0; k < 3; ++k) {
                            g[k] -= u * dr[k];
                            g[3 + k] += u * dN[k];
                            dq[k] += u * dr[k];
                        }
                        for (int k = 0; k < 6; ++k)
                            g[6 + k] += u * dT[k];
                    }
                    continue;
                }
                if (ch->reference >> 48) {
                    ADJ_LEAF(ch->reference);
                    continue;
                }
                if (next != GWN_ORACLE_INVALID_INDEX) {
                    if (sp >= 1024)
                        return NAN;
                    stack[sp++] = next;
                }
                next = (uint32_t)(ch->reference & GWN_ORACLE_OFFSET_MASK);
            }
            if (next != GWN_ORACLE_INVALID_INDEX) {
                node = next;
                continue;
            }
            if (sp == 0)
                break;
            node = stack[--sp];
        }
    }
#undef ADJ_LEAF
    if (grad)
        for (int k = 0; k < 3; ++k)
            dqrow[k] = dq[k];
    return w * inv4pi;
}

static void *adj_thread(void *arg)
{
    adj_job *job = (adj_job *)arg;
    const adj_ctx *c = job->c;
    int grad = c->dq != NULL;
    double value = 0.0;
    for (uint64_t q = job->begin; q < job->end; ++q) {
        double w = adj_query(c, q, job->dslot, job->dtri, grad ? c->dq + 3 * q : NULL, grad);
        value += c->u[q] * w;
    }
    job->value = value;
    return NULL;
}

/* Runs the replay over all queries on `threads` threads; per-thread slot/record buffers summed. */
static void adj_run(adj_ctx *c, double *dslot, double *dtri)
{
    int threads = c->threads < 1 ? 1 : c->threads;
    uint64_t Q = (uint64_t)c->value; /* query count passed through value */
    c->value = 0.0;
    if ((uint64_t)threads > Q)
        threads = Q ? (int)Q : 1;
    size_t ns = 12 * 4 * (size_t)(c->node_count ? c->node_count : 1), nr = 9 * (size_t)(c->nt ? c->nt : 1);
    adj_job *jobs = (adj_job *)calloc((size_t)threads, sizeof(adj_job));
    pthread_t *tid = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    for (int t = 0; t < threads; ++t) {
        jobs[t].c = c;
        jobs[t].begin = Q * (uint64_t)t / (uint64_t)threads;
        jobs[t].end = Q * (uint64_t)(t + 1) / (uint64_t)threads;
        jobs[t].dslot = dslot ? (double *)calloc(ns, sizeof(double)) : NULL;
        jobs[t].dtri = dtri ? (double *)calloc(nr, sizeof(double)) : NULL;
        pthread_create(&tid[t], NULL, adj_thread, &jobs[t]);
    }
    for (int t = 0; t < threads; ++t) {
        pthread_join(tid[t], NULL);
        c->value += jobs[t].value;
        if (dslot)
            for (size_t i = 0; i < ns; ++i)
                dslot[i] += jobs[t].dslot[i];
        if (dtri)
            for (size_t i = 0; i < nr; ++i)
                dtri[i] += jobs[t].dtri[i];
        free(jobs[t].dslot);
        free(jobs[t].dtri);
    }
    free(jobs);
    free(tid);
}

/* Payload adjoint of one slot from its field adjoint (dP, dN, dT). */
static void adj_fields_to_payload(const adj_payload *p, const float *bounds, const double *g, int order,
                                  adj_payload *out)
{
    double P[3], N[3], T[6];
    adj_slot_fields(p, bounds, P, N, T);
    double dP[3] = {g[0], g[1], g[2]}, dN[3] = {g[3], g[4], g[5]};
    memset(out, 0, sizeof(*out));
    if (order >= 1) {
        const double *dT = g + 6;
        /* T0 = n1[0] - Px Nx, T1 = n1[4] - Py Ny, T2 = n1[8] - Pz Nz */
        out->n1[0] += dT[0]; dP[0] -= dT[0] * N[0]; dN[0] -= dT[0] * P[0];
        out->n1[4] += dT[1]; dP[1] -= dT[1] * N[1]; dN[1] -= dT[1] * P[1];
        out->n1[8] += dT[2]; dP[2] -= dT[2] * N[2]; dN[2] -= dT[2] * P[2];
        /* T3 = n1[1] - Py Nx + n1[3] - Px Ny */
        out->n1[1] += dT[3]; out->n1[3] += dT[3];
        dP[1] -= dT[3] * N[0]; dN[0] -= dT[3] * P[1]; dP[0] -= dT[3] * N[1]; dN[1] -= dT[3] * P[0];
        /* T4 = n1[5] - Pz Ny + n1[7] - Py Nz */
        out->n1[5] += dT[4]; out->n1[7] += dT[4];
        dP[2] -= dT[4] * N[1]; dN[1] -= dT[4] * P[2]; dP[1] -= dT[4] * N[2]; dN[2] -= dT[4] * P[1];
        /* T5 = n1[6] - Px Nz + n1[2] - Pz Nx */
        out->n1[6] += dT[5]; out->n1[2] += dT[5];
        dP[0] -= dT[5] * N[2]; dN[2] -= dT[5] * P[0]; dP[2] -= dT[5] * N[0]; dN[0] -= dT[5] * P[2];
    }
    for (int k = 0; k < 3; ++k)
        out->n[k] = dN[k];
    if (p->area > 0.0) {
        /* P = AP / A */
        for (int k = 0; k < 3; ++k) {
            out->ap[k] = dP[k] / p->area;
            out->area -= dP[k] * P[k] / p->area;
        }
    }
}

/* Vertex derivatives of one triangle's payload adjoint. */
static void adj_payload_to_vertices(const double *a, const double *b, const double *c, const adj_payload *g,
                                    double *ga, double *gb, double *gc)
{
    double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double n[3];
    cross3d(e1, e2, n);
    for (int k = 0; k < 3; ++k)
        n[k] *= 0.5;
    double cen[3] = {(a[0] + b[0] + c[0]) / 3.0, (a[1] + b[1] + c[1]) / 3.0, (a[2] + b[2] + c[2]) / 3.0};
    double area = sqrt(dot3d(n, n));
    double gN[3], gcen[3];
    for (int k = 0; k < 3; ++k) {
        gN[k] = g->n[k];
        gcen[k] = g->ap[k] * area;
    }
    double darea = g->area + dot3d(g->ap, cen);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            gN[i] += g->n1[3 * i + j] * cen[j];
            gcen[j] += g->n1[3 * i + j] * n[i];
        }
    if (area > 0.0)
        for (int k = 0; k < 3; ++k)
            gN[k] += darea * n[k] / area;
    /* N = (a x b + b x c + c x a) / 2: d(gN . N)/da = (b - c) x gN / 2, cyclic. */
    double bc[3] = {b[0] - c[0], b[1] - c[1], b[2] - c[2]};
    double ca[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double ab[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    double ta[3], tb[3], tc[3];
    cross3d(bc, gN, ta);
    cross3d(ca, gN, tb);
    cross3d(ab, gN, tc);
    for (int k = 0; k < 3; ++k) {
        ga[k] = 0.5 * ta[k] + gcen[k] / 3.0;
        gb[k] = 0.5 * tb[k] + gcen[k] / 3.0;
        gc[k] = 0.5 * tc[k] + gcen[k] / 3.0;
    }
}

static void adj_push_down(const gwn_oracle_node4_f32 *nodes, uint64_t node, const adj_payload *acc_in,
                          const adj_payload *dpay, adj_payload *dtri_pay, uint64_t nt)
{
    for (int s = 0; s < 4; ++s) {
        uint64_t ref = nodes[node].children[s].reference;
        if (!(ref & GWN_ORACLE_VALID_BIT))
            continue;
        adj_payload acc = *acc_in;
        adj_payload_add(&acc, &dpay[4 * node + s]);
        uint64_t off = ref & GWN_ORACLE_OFFSET_MASK;
        uint32_t cnt = (uint32_t)(ref >> 48);
        if (cnt == 0) {
            adj_push_down(nodes, off, &acc, dpay, dtri_pay, nt);
            continue;
        }
        for (uint32_t k = 0; k < cnt; ++k)
            if (off + k < nt)
                adj_payload_add(&dtri_pay[off + k], &acc);
    }
}

static void adj_mesh_from(adj_mesh *m, double **store, const double *vd, uint64_t V, const uint32_t *prim,
                          const uint32_t *i0, const uint32_t *i1, const uint32_t *i2)
{
    (void)store;
    m->vx = vd;
    m->vy = vd + V;
    m->vz = vd + 2 * V;
    m->prim = prim;
    m->i0 = i0;
    m->i1 = i1;
    m->i2 = i2;
}

double gwn_oracle_winding_taylor_frozen_f64(
    const gwn_oracle_node4_f32 *nodes, uint64_t node_count, gwn_oracle_child_f32 root, const float *triangles,
    uint64_t triangle_count, const float *moments, int order, float beta, const uint32_t *primitive_indices,
    const uint32_t *i0, const uint32_t *i1, const uint32_t *i2, const double *vertices, uint64_t vertex_count,
    const float *qx, const float *qy, const float *qz, const double *u, uint64_t query_count, int threads)
{
    if (order < 0 || order > 1)
        return NAN;
    adj_mesh m;
    adj_mesh_from(&m, NULL, vertices, vertex_count, primitive_indices, i0, i1, i2);
    adj_payload *slot = (adj_payload *)calloc(4 * (node_count ? node_count : 1), sizeof(adj_payload));
    adj_slot_payloads(nodes, node_count, &m, slot);
    adj_ctx c = {nodes, node_count, root, triangles, triangle_count, moments, order,
                 gwn_oracle_moment_field_count(order), beta, &m, slot, qx, qy, qz, u, threads,
                 NULL, NULL, NULL, (double)query_count};
    adj_run(&c, NULL, NULL);
    free(slot);
    return c.value;
}

int gwn_oracle_winding_taylor_adjoint_f64(
    const gwn_oracle_node4_f32 *nodes, uint64_t node_count, gwn_oracle_child_f32 root, const float *triangles,
    uint64_t triangle_count, const float *moments, int order, float beta, const uint32_t *primitive_indices,
    const uint32_t *i0, const uint32_t *i1, const uint32_t *i2, const double *vertices, uint64_t vertex_count,
    const float *qx, const float *qy, const float *qz, const double *u, uint64_t query_count, double *d_vertices,
    double *d_queries, int threads)
{
    if (order < 0 || order > 1)
        return GWN_ORACLE_INVALID_ARGUMENT;
    adj_mesh m;
    adj_mesh_from(&m, NULL, vertices, vertex_count, primitive_indices, i0, i1, i2);
    size_t NS = 4 * (size_t)(node_count ? node_count : 1);
    adj_payload *slot = (adj_payload *)calloc(NS, sizeof(adj_payload));
    adj_slot_payloads(nodes, node_count, &m, slot);
    double *dslot = (double *)calloc(12 * NS, sizeof(double));
    double *dtri = (double *)calloc(9 * (size_t)(triangle_count ? triangle_count : 1), sizeof(double));
    adj_ctx c = {nodes, node_count, root, triangles, triangle_count, moments, order,
                 gwn_oracle_moment_field_count(order), beta, &m, slot, qx, qy, qz, u, threads,
                 NULL, NULL, d_queries, (double)query_count};
    adj_run(&c, dslot, dtri);
    /* Field adjoints -> payload adjoints -> every triangle below. */
    adj_payload *dpay = (adj_payload *)calloc(NS, sizeof(adj_payload));
    adj_payload *dtri_pay = (adj_payload *)calloc(triangle_count ? triangle_count : 1, sizeof(adj_payload));
    for (uint64_t n = 0; n < node_count; ++n)
        for (int s = 0; s < 4; ++s)
            if (nodes[n].children[s].reference & GWN_ORACLE_VALID_BIT)
                adj_fields_to_payload(&slot[4 * n + s], nodes[n].children[s].bounds, dslot + 12 * (4 * n + s), order,
                                      &dpay[4 * n + s]);
    int internal_root = (root.reference & GWN_ORACLE_VALID_BIT) && (root.reference >> 48) == 0 && node_count > 0;
    if (internal_root) {
        adj_payload zero;
        memset(&zero, 0, sizeof(zero));
        adj_push_down(nodes, root.reference & GWN_ORACLE_OFFSET_MASK, &zero, dpay, dtri_pay, triangle_count);
    }
    memset(d_vertices, 0, sizeof(double) * 3 * vertex_count);
    for (uint64_t k = 0; k < triangle_count; ++k) {
        double a[3], b[3], cc[3], ga[3], gb[3], gc[3];
        uint32_t ids[3];
        adj_record_vertices(&m, k, a, b, cc, ids);
        adj_payload_to_vertices(a, b, cc, &dtri_pay[k], ga, gb, gc);
        const double *nt9 = dtri + 9 * k;
        const double *g[3] = {ga, gb, gc};
        for (int t = 0; t < 3; ++t)
            for (int d = 0; d < 3; ++d)
                d_vertices[(size_t)d * vertex_count + ids[t]] += g[t][d] + nt9[3 * t + d];
    }
    free(slot); free(dslot); free(dtri); free(dpay); free(dtri_pay);
    return GWN_ORACLE_OK;
}

/* Magnitude sums of the winding adjoint's terms, for FP32 error envelopes: over every triangle t
 * and query q (no BVH), mag_dv[k][v] = sum |u_q| |d(W_t(q)/4pi)/dv_k| over the triangles t at v and
 * mag_dq[q][k] = |u_q| sum_t |d(W_t(q)/4pi)/dq_k|.  A sum of FP32 terms of these sizes carries a
 * rounding error of a few FP32 ulps of the magnitude sum however much the terms cancel. */
typedef struct {
    const double *vx, *vy, *vz;
    const uint32_t *i0, *i1, *i2;
    uint64_t nt;
    const float *qx, *qy, *qz;
    const double *u;
    uint64_t nq;
    double *tri9;   /* per triangle: 3 vertices x 3 components */
    double *dq;     /* per query: 3 components */
} adj_mag_ctx;

static void adj_mag_vertices(const adj_mag_ctx *c, uint64_t t, double v[3][3])
{
    uint32_t ids[3] = {c->i0[t], c->i1[t], c->i2[t]};
    for (int k = 0; k < 3; ++k) {
        v[k][0] = c->vx[ids[k]];
        v[k][1] = c->vy[ids[k]];
        v[k][2] = c->vz[ids[k]];
    }
}

static void adj_mag_tri_range(void *vctx, uint64_t begin, uint64_t end)
{
    const adj_mag_ctx *c = (const adj_mag_ctx *)vctx;
    double const inv4pi = 1.0 / (4.0 * 3.141592653589793238462643383279502884);
    for (uint64_t t = begin; t < end; ++t) {
        double v[3][3], d[3][3];
        adj_mag_vertices(c, t, v);
        double *o = c->tri9 + 9 * t;
        for (uint64_t q = 0; q < c->nq; ++q) {
            double p[3] = {c->qx[q], c->qy[q], c->qz[q]};
            adj_solid_angle(v[0], v[1], v[2], p, d[0], d[1], d[2]);
            double w = fabs(c->u[q]) * inv4pi;
            for (int i = 0; i < 3; ++i)
                for (int k = 0; k < 3; ++k)
                    o[3 * i + k] += w * fabs(d[i][k]);
        }
    }
}

static void adj_mag_query_range(void *vctx, uint64_t begin, uint64_t end)
{
    const adj_mag_ctx *c = (const adj_mag_ctx *)vctx;
    double const inv4pi = 1.0 / (4.0 * 3.141592653589793238462643383279502884);
    for (uint64_t q = begin; q < end; ++q) {
        double p[3] = {c->qx[q], c->qy[q], c->qz[q]};
        double acc[3] = {0.0, 0.0, 0.0};
        for (uint64_t t = 0; t < c->nt; ++t) {
            double v[3][3], d[3][3];
            adj_mag_vertices(c, t, v);
            adj_solid_angle(v[0], v[1], v[2], p, d[0], d[1], d[2]);
            /* Translation invariance: dW/dq = -(dW/da + dW/db + dW/dc). */
            for (int k = 0; k < 3; ++k)
                acc[k] += fabs(d[0][k] + d[1][k] + d[2][k]);
        }
        for (int k = 0; k < 3; ++k)
            c->dq[3 * q + k] = fabs(c->u[q]) * inv4pi * acc[k];
    }
}

void gwn_oracle_adjoint_magnitude_f64(
    const double *vertices, uint64_t vertex_count, const uint32_t *i0, const uint32_t *i1, const uint32_t *i2,
    uint64_t triangle_count, const float *qx, const float *qy, const float *qz, const double *u, uint64_t count,
    double *mag_dv, double *mag_dq, int threads)
{
    double *tri9 = (double *)calloc(9 * (size_t)(triangle_count ? triangle_count : 1), sizeof(double));
    adj_mag_ctx c = {vertices, vertices + vertex_count, vertices + 2 * vertex_count, i0, i1, i2, triangle_count,
                     qx, qy, qz, u, count, tri9, mag_dq};
    parallel_ranges(adj_mag_tri_range, &c, triangle_count, threads);
    parallel_ranges(adj_mag_query_range, &c, count, threads);
    memset(mag_dv, 0, sizeof(double) * 3 * vertex_count);
    for (uint64_t t = 0; t < triangle_count; ++t) {
        uint32_t ids[3] = {i0[t], i1[t], i2[t]};
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k)
                mag_dv[(size_t)k * vertex_count + ids[i]] += tri9[9 * t + 3 * i + k];
    }
    free(tri9);
}

/* Distance adjoint by brute force in FP64: each point's closest point c on its nearest triangle,
 * n = (p - c) / |p - c|; d d/dp = n and d d/dv_i = -b_i n with b the barycentrics of c (the
 * closest point moves with the triangle).  Returns the distances in `dist`. */
static void adj_closest(const double *p, const double *a, const double *b, const double *c, double *bary)
{
    /* Ericson, Real-Time Collision Detection 5.1.5, in FP64 with barycentric output. */
    double ab[3], ac[3], ap[3];
    for (int k = 0; k < 3; ++k) {
        ab[k] = b[k] - a[k];
        ac[k] = c[k] - a[k];
        ap[k] = p[k] - a[k];
    }
    double d1 = dot3d(ab, ap), d2 = dot3d(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) { bary[0] = 1; bary[1] = 0; bary[2] = 0; return; }
    double bp[3] = {p[0] - b[0], p[1] - b[1], p[2] - b[2]};
    double d3 = dot3d(ab, bp), d4 = dot3d(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) { bary[0] = 0; bary[1] = 1; bary[2] = 0; return; }
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        double v = d1 / (d1 - d3);
        bary[0] = 1 - v; bary[1] = v; bary[2] = 0; return;
    }
    double cp[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
    double d5 = dot3d(ab, cp), d6 = dot3d(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) { bary[0] = 0; bary[1] = 0; bary[2] = 1; return; }
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        double w = d2 / (d2 - d6);
        bary[0] = 1 - w; bary[1] = 0; bary[2] = w; return;
    }
    double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        bary[0] = 0; bary[1] = 1 - w; bary[2] = w; return;
    }
    double den = 1.0 / (va + vb + vc);
    double v = vb * den, w = vc * den;
    bary[0] = 1 - v - w; bary[1] = v; bary[2] = w;
}

typedef struct {
    const double *vx, *vy, *vz;
    const uint32_t *i0, *i1, *i2;
    uint64_t nt, V;
    const float *px, *py, *pz;
    const double *u;
    double *dist, *dp;
    double *dv;   /* per-thread 3V */
} dist_adj_ctx;

typedef struct {
    const dist_adj_ctx *c;
    uint64_t begin, end;
    double *dv;
} dist_adj_job;

static void *dist_adj_thread(void *arg)
{
    dist_adj_job *job = (dist_adj_job *)arg;
    const dist_adj_ctx *c = job->c;
    for (uint64_t i = job->begin; i < job->end; ++i) {
        double p[3] = {c->px[i], c->py[i], c->pz[i]};
        double best = INFINITY, bb[3] = {0, 0, 0}, bc[3] = {0, 0, 0};
        uint64_t bt = 0;
        for (uint64_t t = 0; t < c->nt; ++t) {
            uint32_t ids[3] = {c->i0[t], c->i1[t], c->i2[t]};
            double v[3][3];
            for (int k = 0; k < 3; ++k) {
                v[k][0] = c->vx[ids[k]];
                v[k][1] = c->vy[ids[k]];
                v[k][2] = c->vz[ids[k]];
            }
            double bary[3];
            adj_closest(p, v[0], v[1], v[2], bary);
            double cp[3], d2 = 0.0;
            for (int k = 0; k < 3; ++k) {
                cp[k] = bary[0] * v[0][k] + bary[1] * v[1][k] + bary[2] * v[2][k];
                d2 += (p[k] - cp[k]) * (p[k] - cp[k]);
            }
            if (d2 < best) {
                best = d2;
                bt = t;
                memcpy(bb, bary, sizeof(bb));
                memcpy(bc, cp, sizeof(bc));
            }
        }
        double d = sqrt(best);
        c->dist[i] = d;
        double n[3] = {0, 0, 0};
        if (d > 0.0)
            for (int k = 0; k < 3; ++k)
                n[k] = (p[k] - bc[k]) / d;
        for (int k = 0; k < 3; ++k)
            c->dp[3 * i + k] = c->u[i] * n[k];
        uint32_t ids[3] = {c->i0[bt], c->i1[bt], c->i2[bt]};
        for (int t = 0; t < 3; ++t)
            for (int k = 0; k < 3; ++k)
                job->dv[(size_t)k * c->V + ids[t]] -= c->u[i] * bb[t] * n[k];
    }
    return NULL;
}

void gwn_oracle_distance_adjoint_f64(
    const double *vertices, uint64_t vertex_count, const uint32_t *i0, const uint32_t *i1, const uint32_t *i2,
    uint64_t triangle_count, const float *px, const float *py, const float *pz, const double *u, uint64_t count,
    double *dist, double *d_vertices, double *d_points, int threads)
{
    dist_adj_ctx c = {vertices, vertices + vertex_count, vertices + 2 * vertex_count, i0, i1, i2, triangle_count,
                      vertex_count, px, py, pz, u, dist, d_points, NULL};
    if (threads < 1)
        threads = 1;
    if ((uint64_t)threads > count)
        threads = count ? (int)count : 1;
    dist_adj_job *jobs = (dist_adj_job *)calloc((size_t)threads, sizeof(dist_adj_job));
    pthread_t *tid = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    for (int t = 0; t < threads; ++t) {
        jobs[t].c = &c;
        jobs[t].begin = count * (uint64_t)t / (uint64_t)threads;
        jobs[t].end = count * (uint64_t)(t + 1) / (uint64_t)threads;
        jobs[t].dv = (double *)calloc(3 * (size_t)(vertex_count ? vertex_count : 1), sizeof(double));
        pthread_create(&tid[t], NULL, dist_adj_thread, &jobs[t]);
    }
    memset(d_vertices, 0, sizeof(double) * 3 * vertex_count);
    for (int t = 0; t < threads; ++t) {
        pthread_join(tid[t], NULL);
        for (size_t i = 0; i < 3 * (size_t)vertex_count; ++i)
            d_vertices[i] += jobs[t].dv[i];
        free(jobs[t].dv);
    }
    free(jobs);
    free(tid);
}
