"""bench.py's use of the committed counter summaries: their figures enter the bench line only
while the traversal kernel's device source is the one they were measured at."""

import bench


def test_summaries_match_the_current_traversal_source():
    """The committed summaries were measured on the kernel the tree now holds (a kernel change
    that is not re-profiled shows up here first)."""
    assert bench.traffic_profile().get("current") is True
    assert bench.valu_profile().get("current") is True
    assert bench.traffic_profile()["hbm_bytes_per_launch"] > 0
    assert 0.0 < bench.valu_profile()["valu_lane_utilization"] <= 1.0


def test_figures_are_left_out_when_the_source_differs(monkeypatch):
    monkeypatch.setattr(bench, "traversal_source_digest", lambda: "0" * 64)
    t, v = bench.traffic_profile(), bench.valu_profile()
    assert t["current"] is False and "hbm_bytes_per_launch" not in t
    assert v["current"] is False and "valu_lane_utilization" not in v


def test_digest_covers_the_traversal_text(monkeypatch, tmp_path):
    """An edit inside the hashed region changes the digest; one outside it does not."""
    import os
    import shutil

    for rel in ("include/gwn/detail/gwn_device_layout.hpp", "include/gwn/detail/gwn_device_winding.hpp",
                "smallgwn_amd/csrc/gwn_winding.hip"):
        os.makedirs(tmp_path / os.path.dirname(rel), exist_ok=True)
        shutil.copy(os.path.join(bench.REPO, rel), tmp_path / rel)
    monkeypatch.setattr(bench, "REPO", str(tmp_path))
    base = bench.traversal_source_digest()
    src = tmp_path / "smallgwn_amd/csrc/gwn_winding.hip"
    text = src.read_text()
    src.write_text(text + "\n// after the hashed region\n")
    assert bench.traversal_source_digest() == base
    marker = "bool const in = (mask & my_bit) != 0ull;"
    assert marker in text
    src.write_text(text.replace(marker, marker + " // edited", 1))
    assert bench.traversal_source_digest() != base
