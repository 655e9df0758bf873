"""The scop front end: parses C statements into tap tables, and the committed
generated header is what it produces."""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
spec = importlib.util.spec_from_file_location(
    "scop2desc", os.path.join(ROOT, "an5d-artifact_amd", "frontend", "scop2desc.py"))
S = importlib.util.module_from_spec(spec)
sys.modules["scop2desc"] = S
spec.loader.exec_module(S)

TEMPLATE = """#define BENCH_DIM {dim}
#define BENCH_FPP 3
#define BENCH_RAD 1
#pragma scop
    for (int t = 0; t < timestep; t++)
      for (int i = 1; i < n; i++)
          A[(t+1)%2]{lhs} = {rhs};
#pragma endscop
"""


def parse(tmp_path, rhs, dim=2):
    lhs = "".join(f"[{v}]" for v in "ijk"[:dim])
    f = tmp_path / "x.c"
    f.write_text(TEMPLATE.format(dim=dim, lhs=lhs, rhs=rhs))
    return S.parse_source(str(f))


def test_linear_chain_with_signs_and_divisor(tmp_path):
    st = parse(tmp_path, "(2.0f * A[t%2][i-1][j] - 0.5f*A[t%2][i][j+1] + A[t%2][i][j] * 3.f) / 7")
    assert st.kind == "linear" and st.divisor == 7
    assert [(t.sign, t.coef, t.off) for t in st.taps] == [
        (1, "2.0f", (-1, 0, 0)), (-1, "0.5f", (0, 0, 1)), (1, "3.f", (0, 0, 0))]


def test_negative_literal_and_3d_offsets(tmp_path):
    st = parse(tmp_path, "-3.240f*A[t%2][i-4][j][k] + 0.5f*A[t%2][i][j+2][k-1]", dim=3)
    assert [(t.sign, t.coef, t.off) for t in st.taps] == [(1, "-3.240f", (-4, 0, 0)), (1, "0.5f", (0, 2, -1))]


def test_nonlinear_is_flagged(tmp_path):
    st = parse(tmp_path, "A[t%2][i][j] + 1.0f / sqrt(0.1f + (A[t%2][i][j]-A[t%2][i-1][j]))")
    assert st.kind == "nonlinear"
    assert (0, 0, 0) in st.points and (-1, 0, 0) in st.points


def test_rejects_wrong_time_index(tmp_path):
    with pytest.raises(S.ParseError):
        parse(tmp_path, "A[t%3][i][j]")


def test_generated_header_is_current():
    ref = "/root/reference"
    if not os.path.exists(os.path.join(ref, "j2d5pt.c")):
        pytest.skip("reference sources not present")
    sts = [S.parse_source(os.path.join(ref, n + ".c")) for n in S.STENCILS]
    text = S.render_header(sts)
    committed = open(os.path.join(ROOT, "an5d-artifact_amd", "csrc", "stencils_gen.h")).read()
    assert text == committed
