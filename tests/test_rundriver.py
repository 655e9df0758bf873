"""The run driver writes the reference's run log format
(compiled/run.sh:13,27-35; summarize.sh:44-49) from the binaries' output."""
import csv
import os
import stat

import pytest

from pyan5d import rundriver

CAPTURED = """[Device warm-up]
1.2345 ms

[Device execution]
      0: 12345.6789 GFLOPS,    10.0000 ms
      1: 12400.0000 GFLOPS,     9.9558 ms
Average: 12372.7928 GFLOPS,     9.9779 ms
GStencil/s:  1237.2793
"""


def _fake_bin(tmp_path, name, text, rc=0):
    p = tmp_path / name
    p.write_text("#!/bin/sh\ncat <<'EOF'\n" + text + "EOF\n" + f"exit {rc}\n")
    p.chmod(p.stat().st_mode | stat.S_IEXEC)
    return p


def test_measured_field_is_run_sh_field_two():
    assert rundriver.measured_gflops(CAPTURED) == "12372.7928"
    assert rundriver.measured_gflops("Average:    1.5000 GFLOPS,  3.0 ms\n") == "1.5000"
    assert rundriver.measured_gflops("HIP error: out of memory\n") is None


def test_model_ranked_configurations():
    cfgs = rundriver.top_configs("j2d5pt", "float", 16384, k=5)
    assert len(cfgs) == 5
    assert all(c["bS2"] is None and c["bS1"] % 64 == 0 for c in cfgs)
    assert [c["model"] for c in cfgs] == sorted((c["model"] for c in cfgs), reverse=True)
    c3 = rundriver.top_configs("star3d1r", "double", 512, k=3)
    assert all(c["bS2"] is not None for c in c3)


def test_csv_from_runs_and_summary(tmp_path):
    _fake_bin(tmp_path, "j2d5pt.float", CAPTURED)
    _fake_bin(tmp_path, "star3d1r.float", "[Device warm-up]\nHIP error: no device\n", rc=134)
    cfgs = [dict(stencil="j2d5pt", variant=0, bt=7, sl=64, bS1=256, bS2=None, model=21000.4),
            dict(stencil="star3d1r", variant=3, bt=3, sl=128, bS1=192, bS2=32, model=15000.0)]
    rows = [rundriver.csv_row(c, rundriver.run_config(c, "float", 64, 10, 1, bin_dir=str(tmp_path),
                                                      log_dir=str(tmp_path / "logs")))
            for c in cfgs]
    out = tmp_path / "float_run.csv"
    rundriver.write_csv(str(out), rows)
    text = out.read_text()
    assert text.splitlines() == ["Stencil,bS1,bS2,bT,sl,Model,Measured",
                                 "j2d5pt,256,,7,64,21000,12372.7928",
                                 "star3d1r,192,32,3,128,15000,Failed"]
    assert rundriver.summarize(str(out)) == ["j2d5pt,12372.7928", "star3d1r,Failed"]
    assert os.path.exists(tmp_path / "logs" / "j2d5pt-v0-7-64.log")


def test_reference_list_input(tmp_path):
    lst = tmp_path / "v100_float.csv"
    lst.write_text('j2d5pt,256,,10,256,"8,144"\nstar3d1r,32,32,5,128,"3,886"\n')
    rows = rundriver.read_reference_list(str(lst))
    assert rows[0]["bt"] == 10 and rows[0]["sl"] == 256 and rows[0]["model"] == 8144.0
    assert rows[1]["bS2"] == "32" and rows[1]["variant"] == -1


@pytest.mark.gpu
def test_driver_runs_the_binaries(tmp_path):
    cfgs = rundriver.top_configs("j2d5pt", "float", 1024, k=2)
    print(f"case rundriver {cfgs}", flush=True)
    rows = [rundriver.csv_row(c, rundriver.run_config(c, "float", 1024, 20, 1)) for c in cfgs]
    for r in rows:
        assert float(r[6]) > 0, r
