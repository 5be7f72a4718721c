// Compiles the reference-named C++ surface (include/gwn/gwn.hpp) against libgwn_hip.so and runs
// the calls whose outcome is decided on the host: object lifetimes and the validation errors the
// reference tests assert (tests/unit_gwn_query_stack.cu:71-76 and the empty-object contracts of
// tests/unit_gwn_query_gradient.cu:280-300).  No GPU work is enqueued.
#include <gwn/gwn.hpp>

#include <cstdint>
#include <cstdio>
#include <string>

namespace {
int failures = 0;
void expect(bool ok, char const *what) {
    if (!ok) {
        std::printf("FAIL: %s\n", what);
        ++failures;
    }
}
} // namespace

int main() {
    using Real = float;
    using Index = std::uint32_t;
    gwn::gwn_geometry_object<Real, Index> geometry;
    gwn::gwn_bvh4_object<Real, Index> bvh;
    gwn::gwn_bvh4_moment_object<1, Real, Index> moment;
    expect(geometry.vertex_count() == 0 && geometry.triangle_count() == 0, "empty geometry counts");
    expect(!bvh.has_data(), "empty BVH has no data");

    // Move-only ownership.
    gwn::gwn_bvh4_object<Real, Index> moved(std::move(bvh));
    expect(!moved.has_data(), "moved BVH is still empty");

    // H-PLOC radius validation happens before any device work.
    gwn::gwn_bvh_build_options opts{};
    opts.hploc_search_radius = 0;
    gwn::gwn_status st = gwn::gwn_build_bvh(geometry, moved, opts);
    expect(st.error() == gwn::gwn_error::invalid_argument, "H-PLOC radius 0 rejected");

    // Queries on an empty BVH are invalid arguments, not crashes.
    gwn::gwn_device_span<Real const> const empty_in{};
    gwn::gwn_device_span<Real> const empty_out{};
    st = gwn::gwn_compute_winding_number_exact_batch(moved, empty_in, empty_in, empty_in, empty_out);
    expect(st.error() == gwn::gwn_error::invalid_argument, "exact batch on empty BVH");
    st = gwn::gwn_compute_winding_number_taylor_batch<1>(moved, moment, empty_in, empty_in, empty_in, empty_out);
    expect(st.error() == gwn::gwn_error::invalid_argument, "Taylor batch on empty BVH");
    expect(st.message().find("BVH") != std::string::npos, "Taylor batch message names the BVH");

    st = gwn::gwn_compute_winding_gradient_taylor_batch<1>(moved, moment, empty_in, empty_in, empty_in, empty_out,
                                                          empty_out, empty_out);
    expect(st.error() == gwn::gwn_error::invalid_argument, "gradient batch on empty BVH");

    // Length mismatch is caught by the header before the call.
    Real dummy[4] = {0, 0, 0, 0};
    gwn::gwn_device_span<Real const> const four(dummy, 4);
    gwn::gwn_device_span<Real> const three(dummy, 3);
    st = gwn::gwn_compute_winding_number_exact_batch(moved, four, four, four, three);
    expect(st.error() == gwn::gwn_error::invalid_argument, "span length mismatch");

    // Refit of an empty BVH.
    st = gwn::gwn_refit_bvh(geometry, moved);
    expect(st.error() == gwn::gwn_error::invalid_argument, "refit of empty BVH");
    st = gwn::gwn_refit_bvh_moment(moved, moment);
    expect(st.error() == gwn::gwn_error::invalid_argument, "moment refit of empty BVH");

    // Boundary chain and antipodal batches (tests/unit_gwn_query_antipodal.cu:577-640): an
    // unbuilt chain and an empty geometry are rejected before any launch.
    gwn::gwn_boundary_chain_object<Index> chain;
    expect(!chain.is_built() && chain.edge_count() == 0, "new chain is unbuilt and empty");
    st = gwn::gwn_compute_winding_number_antipodal_batch(geometry, moved, chain, empty_in, empty_in, empty_in,
                                                         empty_out);
    expect(st.error() == gwn::gwn_error::invalid_argument, "antipodal batch on empty BVH");
    st = gwn::gwn_compute_winding_gradient_antipodal_batch(geometry, chain, empty_in, empty_in, empty_in, empty_out,
                                                           empty_out, empty_out);
    expect(st.error() == gwn::gwn_error::invalid_argument, "antipodal gradient with an unbuilt chain");
    expect(st.message().find("unbuilt") != std::string::npos, "antipodal gradient names the unbuilt chain");

    // Stream mixin (tests/unit_gwn_stream.cu:7-20): explicit bindings survive set_stream and swap.
    {
        gwn::gwn_bvh4_object<Real, Index> first;
        gwn::gwn_bvh4_object<Real, Index> second;
        expect(first.stream() == gwn::gwn_default_stream(), "new object is bound to the default stream");
        auto const sa = reinterpret_cast<hipStream_t>(std::uintptr_t(0xAAAA));
        auto const sb = reinterpret_cast<hipStream_t>(std::uintptr_t(0xBBBB));
        first.set_stream(sa);
        second.set_stream(sb);
        swap(first, second);
        expect(first.stream() == sb && second.stream() == sa, "swap exchanges stream bindings");
        // clear() on an empty object releases nothing and keeps the binding.
        first.clear();
        expect(first.stream() == sb && !first.has_data(), "clear keeps the binding");
        gwn::gwn_bvh4_moment_object<0, Real, Index> m;
        m.set_stream(sa);
        m.clear();
        expect(m.stream() == sa && !m.has_data() && !m.is_valid_for(first), "moment clear keeps the binding");
        gwn::gwn_geometry_object<Real, Index> g;
        g.set_stream(sb);
        expect(g.stream() == sb && g.accessor().x == nullptr && !g.has_data(), "geometry binding and accessor");
        gwn::gwn_boundary_chain_object<Index> c;
        c.set_stream(sa);
        expect(c.stream() == sa && c.accessor().is_built == 0, "boundary binding and accessor");
        // Width 2 and 3 objects exist; a width outside [2, 4] would not compile.
        gwn::gwn_bvh_object<2, Real, Index> b2;
        gwn::gwn_bvh_object<3, Real, Index> b3;
        expect(!b2.has_data() && !b3.has_data() && b2.accessor().nodes == nullptr, "narrow BVH objects");
    }

    // Real = double, Index = uint64 objects (include/gwn/gwn_bvh_build.cuh:34-39): the FP64 workflow's
    // host-side validation.  An index outside [0, vertex_count) is rejected before any copy.
    {
        using R64 = double;
        using I64 = std::uint64_t;
        gwn::gwn_geometry_object<R64, I64> g64;
        gwn::gwn_bvh4_object<R64, I64> b64;
        gwn::gwn_bvh4_moment_object<2, R64, I64> m64;
        R64 const vx[3] = {0, 1, 0}, vy[3] = {0, 0, 1}, vz[3] = {0, 0, 0};
        I64 const bad[1] = {3}, ok1[1] = {1}, ok2[1] = {2};
        gwn::gwn_status s64 = gwn::gwn_upload_geometry(
            g64, gwn::gwn_host_span<R64 const>(vx, 3), gwn::gwn_host_span<R64 const>(vy, 3),
            gwn::gwn_host_span<R64 const>(vz, 3), gwn::gwn_host_span<I64 const>(bad, 1),
            gwn::gwn_host_span<I64 const>(ok1, 1), gwn::gwn_host_span<I64 const>(ok2, 1));
        expect(s64.error() == gwn::gwn_error::invalid_argument, "uint64 index out of range rejected");
        expect(!g64.has_data(), "rejected upload leaves the geometry empty");
        gwn::gwn_device_span<R64 const> const in64{};
        gwn::gwn_device_span<R64> const out64{};
        s64 = gwn::gwn_compute_winding_number_exact_batch(b64, in64, in64, in64, out64);
        expect(s64.error() == gwn::gwn_error::invalid_argument, "double exact batch on empty BVH");
        s64 = gwn::gwn_compute_winding_number_taylor_batch<2>(b64, m64, in64, in64, in64, out64, R64(1e6));
        expect(s64.error() == gwn::gwn_error::invalid_argument, "double Taylor batch on empty BVH");
        s64 = gwn::gwn_refit_bvh(g64, b64);
        expect(s64.error() == gwn::gwn_error::invalid_argument, "double refit of empty BVH");
        s64 = gwn::gwn_refit_bvh_moment(b64, m64);
        expect(s64.error() == gwn::gwn_error::invalid_argument, "double moment refit of empty BVH");
        expect(!m64.has_data() && !m64.is_valid_for(b64), "empty double moment");
    }

    if (failures == 0)
        std::printf("gwn.hpp surface: all checks passed\n");
    return failures == 0 ? 0 : 1;
}
