"""The compiler resource table (an5d-artifact_amd/resources.py): it reads the
gfx950 code-object metadata of this build's objects.  CPU only; skipped where
the build's objects are absent (they are not shipped with the library)."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "an5d-artifact_amd", "build")
sys.path.insert(0, os.path.join(ROOT, "an5d-artifact_amd"))

pytestmark = pytest.mark.skipif(not os.path.isdir(BUILD) or not os.path.exists(
    os.path.join(BUILD, "fwd_12_0.o")), reason="no build objects here")


def test_resource_table_covers_the_kernels(tmp_path):
    import resources
    rows = resources.collect(BUILD)
    fams = {r["family"] for r in rows}
    assert {"sweep_kernel", "sweep_kernel_cap", "rows_kernel", "plane_kernel", "direct_kernel",
            "adjoint_kernel"} <= fams
    # every instance has a register count and a non-negative scratch size
    assert all(r["vgpr"] > 0 and r["scratch"] >= 0 for r in rows)
    # the headline kernel (star3d1r FP32, bt=5, 128 x 32 tile, shared products) fits without spills
    head = [r for r in rows if r["family"] == "sweep_kernel" and r["stencil"] == "star3d1r" and r["dtype"] == "f32"
            and r["params"].get("bt") == "5" and r["params"].get("ck") == "2" and r["params"].get("nw") == "16"
            and r["params"].get("share") == "1"]
    assert len(head) == 1 and head[0]["scratch"] == 0
    # capped twins are held to their cap
    for r in rows:
        if r["family"] == "sweep_kernel_cap":
            w = int(r["params"]["waves_per_eu"])
            assert r["vgpr"] + r["agpr"] <= 512 // w
    tsv, md = tmp_path / "r.tsv", tmp_path / "r.md"
    resources.write_tsv(rows, str(tsv))
    resources.write_md(rows, str(md), "cmd")
    assert len(open(tsv).read().splitlines()) == len(rows) + 1
    assert "instances with scratch" in open(md).read().lower()
