// Compiles the reference-named C++ surface (include/gwn/gwn.hpp) against libgwn_hip.so and runs
// the calls whose outcome is decided on the host: object lifetimes and the validation errors the
// reference tests assert (tests/unit_gwn_query_stack.cu:71-76 and the empty-object contracts of
// tests/unit_gwn_query_gradient.cu:280-300).  No GPU work is enqueued.
#include <gwn/gwn.hpp>

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

    if (failures == 0)
        std::printf("gwn.hpp surface: all checks passed\n");
    return failures == 0 ? 0 : 1;
}
