"""Stream binding and storage lifetime of the objects, on two non-blocking streams.

Restates the reference's stream contract suite (tests/unit_gwn_stream.cu,
tests/unit_gwn_device_array.cu:18-203, tests/unit_gwn_geometry.cu:202-256,
tests/unit_gwn_boundary.cu:213-269, tests/unit_gwn_bvh_moment.cu:211-268) through the C ABI:

- an object is bound to the stream of the operation that last replaced or wrote its storage;
  set_stream rebinds, clear releases and keeps the binding, and get_accessor snapshots it;
- a replacement made on stream B releases the old storage on the old binding A, behind readers
  already queued there.  The readers wait behind a bounded device sleep on A while B replaces the
  object and then allocates storage of the same sizes; had the old storage been released on B,
  those allocations would reuse it and the readers would see their bytes.
"""

import numpy as np
import pytest

from smallgwn_amd import synth

pytestmark = pytest.mark.gpu

SLEEP_CYCLES = 50_000_000  # bounded spin on stream A (tens of milliseconds)


def _streams():
    import torch

    return torch.cuda.Stream(), torch.cuda.Stream()


def _upload(gwn, mesh, stream):
    geom = gwn.GeometryObject()
    gwn.gwn_upload_geometry(geom, mesh.vx, mesh.vy, mesh.vz, mesh.i0, mesh.i1, mesh.i2, stream=stream).raise_if_error()
    return geom


def test_bindings_follow_operations_and_clear_keeps_them(gwn, dev):
    import torch

    a, b = _streams()
    mesh = synth.open_cube()
    geom = _upload(gwn, mesh, a)
    assert geom.stream() == a.cuda_stream
    acc = geom.accessor()
    assert (acc["vertex_count"], acc["triangle_count"]) == (8, 10) and acc["x"] and acc["i2"]
    bvh = gwn.BvhObject()
    gwn.gwn_build_bvh(geom, bvh, stream=a).raise_if_error()
    assert bvh.stream() == a.cuda_stream
    acc = bvh.accessor()
    assert acc["revision"] == bvh.info()["revision"] != 0 and acc["nodes"] and acc["triangles"]
    mom = gwn.MomentObject(0)
    gwn.gwn_refit_bvh_moment(bvh, mom, stream=a).raise_if_error()
    macc = mom.accessor()
    assert macc["bvh_revision"] == acc["revision"] and macc["record_bytes"] == 128 and macc["records"]
    chain = gwn.BoundaryChainObject()
    gwn.gwn_build_boundary_chain(geom, chain, stream=a).raise_if_error()
    cacc = chain.accessor()
    assert cacc["edge_count"] == 4 and cacc["is_built"] == 1

    # Writes in place rebind (gwn_update_geometry, gwn_refit_bvh): the storage now holds B's work.
    gwn.gwn_update_geometry(geom, mesh.vx * 2, mesh.vy, mesh.vz, stream=b).raise_if_error()
    assert geom.stream() == b.cuda_stream
    gwn.gwn_refit_bvh(geom, bvh, stream=b).raise_if_error()
    assert bvh.stream() == b.cuda_stream
    b.synchronize()
    assert float(bvh.info()["root_bounds"][3]) == 2.0  # max x of the stretched cube

    # clear_resets_state_before_explicit_rebind (tests/unit_gwn_bvh_moment.cu:211-224).
    assert mom.stream() == a.cuda_stream
    mom.clear()
    mom.set_stream(b)
    assert mom.stream() == b.cuda_stream
    assert mom.accessor()["records"] is None and mom.accessor()["node_count"] == 0
    # A cleared moment is no longer usable with the BVH.
    q = [torch.zeros(1, device=dev) for _ in range(3)]
    out = torch.empty(1, device=dev)
    assert gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, *q, out).error == gwn.Error.INVALID_ARGUMENT

    for obj in (geom, bvh, chain):
        obj.set_stream(a)
        assert obj.stream() == a.cuda_stream
        obj.clear()
        assert obj.stream() == a.cuda_stream and not obj.has_data()
    assert geom.counts() == (0, 0) and bvh.info()["revision"] == 0 and chain.info()["edge_count"] == 0
    # A cleared geometry is an empty one: it builds an empty, unqueryable BVH.
    assert gwn.gwn_build_bvh(geom, bvh).is_ok()
    assert not bvh.has_data()
    torch.cuda.synchronize()


def _churn(gwn, mesh, stream, count=6):
    """Allocations of the replaced storage's sizes on the replacement stream."""
    keep = []
    for k in range(count):
        g = gwn.GeometryObject()
        gwn.gwn_upload_geometry(g, mesh.vx + 100.0 + k, mesh.vy - 50.0, mesh.vz * 3.0, mesh.i2, mesh.i1, mesh.i0,
                                stream=stream).raise_if_error()
        keep.append(g)
    return keep


def test_geometry_replacement_releases_old_storage_on_its_bound_stream(gwn, dev):
    """tests/unit_gwn_geometry.cu:202-256: a reader of the old positions queued on A still sees
    them after an upload on B replaced the object."""
    import torch

    a, b = _streams()
    mesh = synth.open_cube()
    geom = _upload(gwn, mesh, a)
    chain = gwn.BoundaryChainObject()
    gwn.gwn_build_boundary_chain(geom, chain, stream=a).raise_if_error()
    qs = synth.uniform_box_points(5, 256, (-2.0,) * 3, (2.0,) * 3)
    q = [torch.from_numpy(v).to(dev) for v in qs]
    torch.cuda.synchronize()

    def read(stream):
        g = [torch.empty(256, device=dev) for _ in range(3)]
        gwn.gwn_compute_winding_gradient_antipodal_batch(geom, chain, *q, *g, stream=stream).raise_if_error()
        return g

    first = read(a)
    a.synchronize()  # torch streams are non-blocking: copy out only once A has finished
    expect = [t.cpu().numpy() for t in first]
    with torch.cuda.stream(a):
        torch.cuda._sleep(SLEEP_CYCLES)
    late = read(a)  # queued behind the sleep, on the old binding
    moved = synth._mesh_from_arrays(np.stack([mesh.vx + 7.0, mesh.vy, mesh.vz], 1),
                                    np.stack([mesh.i0, mesh.i1, mesh.i2], 1))
    gwn.gwn_upload_geometry(geom, moved.vx, moved.vy, moved.vz, moved.i0, moved.i1, moved.i2,
                            stream=b).raise_if_error()
    assert geom.stream() == b.cuda_stream
    keep = _churn(gwn, mesh, b)
    b.synchronize()
    a.synchronize()
    for got, want in zip(late, expect):
        assert np.array_equal(got.cpu().numpy(), want)
    del keep


def test_bvh_and_moment_replacement_release_on_the_old_stream(gwn, dev):
    """tests/unit_gwn_bvh_moment.cu:226-268 and the BVH counterpart: Taylor readers of the old
    tree and coefficients queued on A, rebuild and moment refit on B."""
    import torch

    a, b = _streams()
    mesh = synth.displaced_sphere(48, 24)
    geom = _upload(gwn, mesh, a)
    bvh = gwn.BvhObject()
    gwn.gwn_build_bvh(geom, bvh, stream=a).raise_if_error()
    mom = gwn.MomentObject(1)
    gwn.gwn_refit_bvh_moment(bvh, mom, stream=a).raise_if_error()
    n = 1 << 16
    q = [torch.from_numpy(v).to(dev) for v in synth.uniform_box_points(6, n, (-1.5,) * 3, (1.5,) * 3)]
    torch.cuda.synchronize()

    def read(stream):
        out = torch.empty(n, device=dev)
        gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, *q, out, stream=stream).raise_if_error()
        return out

    first = read(a)
    a.synchronize()
    expect = first.cpu().numpy()
    with torch.cuda.stream(a):
        torch.cuda._sleep(SLEEP_CYCLES)
    late = read(a)
    # Replace the moment, then the BVH (radius 2 instead of the default), both on B, then allocate
    # the same sizes on B: trees of the same mesh and their moment objects.
    gwn.gwn_refit_bvh_moment(bvh, mom, stream=b).raise_if_error()
    assert mom.stream() == b.cuda_stream
    gwn.gwn_build_bvh(geom, bvh, gwn.BuildOptions(method=gwn.BuildMethod.HPLOC, hploc_search_radius=2),
                      stream=b).raise_if_error()
    assert bvh.stream() == b.cuda_stream
    keep = []
    for _ in range(3):
        other = gwn.BvhObject()
        gwn.gwn_build_bvh(geom, other, stream=b).raise_if_error()
        m2 = gwn.MomentObject(1)
        gwn.gwn_refit_bvh_moment(other, m2, stream=b).raise_if_error()
        keep += [other, m2]
    b.synchronize()
    a.synchronize()
    assert np.array_equal(late.cpu().numpy(), expect)
    # The stale moment is refused with the rebuilt tree until refit.
    out = torch.empty(n, device=dev)
    assert gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, *q, out).error == gwn.Error.INVALID_ARGUMENT
    del keep


def test_boundary_replacement_releases_old_storage_on_its_bound_stream(gwn, dev):
    """tests/unit_gwn_boundary.cu:213-269: antipodal readers of the old chain queued on A; the
    chain is rebuilt on B from a geometry with different topology."""
    import torch

    a, b = _streams()
    mesh = synth.open_cube()
    geom = _upload(gwn, mesh, a)
    bvh = gwn.BvhObject()
    gwn.gwn_build_bvh(geom, bvh, stream=a).raise_if_error()
    chain = gwn.BoundaryChainObject()
    gwn.gwn_build_boundary_chain(geom, chain, stream=a).raise_if_error()
    qs = synth.uniform_box_points(7, 512, (-2.0,) * 3, (2.0,) * 3)
    q = [torch.from_numpy(v).to(dev) for v in qs]
    torch.cuda.synchronize()

    def read(stream):
        g = [torch.empty(512, device=dev) for _ in range(3)]
        gwn.gwn_compute_winding_gradient_antipodal_batch(geom, chain, *q, *g, stream=stream).raise_if_error()
        return g

    first = read(a)
    a.synchronize()
    expect = [t.cpu().numpy() for t in first]
    with torch.cuda.stream(a):
        torch.cuda._sleep(SLEEP_CYCLES)
    late = read(a)
    flipped = gwn.GeometryObject()
    gwn.gwn_upload_geometry(flipped, mesh.vx, mesh.vy, mesh.vz, mesh.i2, mesh.i1, mesh.i0, stream=b).raise_if_error()
    gwn.gwn_build_boundary_chain(flipped, chain, stream=b).raise_if_error()
    assert chain.stream() == b.cuda_stream
    keep = []
    for _ in range(4):
        c2 = gwn.BoundaryChainObject()
        gwn.gwn_build_boundary_chain(flipped, c2, stream=b).raise_if_error()
        keep.append(c2)
    b.synchronize()
    a.synchronize()
    for got, want in zip(late, expect):
        assert np.array_equal(got.cpu().numpy(), want)
    del keep
