"""Autotuner: the analytic prior (CPU) and the measured search (GPU)."""
import numpy as np
import pytest

import pyan5d


def test_model_prefers_sensible_configs():
    # deeper temporal blocks raise the estimate of a bandwidth-bound 2D star
    g1 = pyan5d.model_estimate("j2d5pt", [16384, 16384], bt=1, variant=0)
    g4 = pyan5d.model_estimate("j2d5pt", [16384, 16384], bt=4, variant=0)
    assert g4 > g1 > 0
    # a grid far too small to fill the GPU is estimated lower per update
    small = pyan5d.model_estimate("j2d5pt", [256, 256], bt=4, variant=0)
    assert small < g4
    # FMA mode costs fewer instructions
    gf = pyan5d.model_estimate("j2d5pt", [16384, 16384], bt=8, variant=0, mode=pyan5d.FAST)
    ge = pyan5d.model_estimate("j2d5pt", [16384, 16384], bt=8, variant=0, mode=pyan5d.EXACT)
    assert gf > ge


def test_tune_table_roundtrip(tmp_path):
    path = str(tmp_path / "t.tsv")
    (tmp_path / "t.tsv").write_text("# header\nj2d5pt 0 0 1000 1000 0 0 3 64 1234.5\nnope 0 0 1 1 0 0 1 0 1\n")
    pyan5d.tune_clear()
    pyan5d.tune_load(path)
    out = str(tmp_path / "u.tsv")
    pyan5d.tune_save(out)
    lines = [l for l in open(out) if not l.startswith("#")]
    # the last column is the chosen kernel's register cap (variant 0 of
    # j2d5pt FP32 has none); it is informational and ignored on load
    assert lines == ["j2d5pt 0 0 1000 1000 0 0 3 64 1234.5 0\n"]
    pyan5d.tune_clear()


def test_tune_table_records_register_cap(tmp_path):
    """The cap column: a capped twin's deep block reports its cap, its
    shallow blocks (which share the uncapped kernels) report none."""
    vs = pyan5d.variants("j2d5pt", pyan5d.FLOAT32)
    capped = [v for v in vs if v["regcap"] > 0]
    assert capped and all(v["kind"] == 0 for v in capped)
    v = capped[0]
    base = [u for u in vs if u["regcap"] == 0 and (u["ck"], u["cj"], u["nw"]) == (v["ck"], v["cj"], v["nw"])]
    assert base and base[0]["btmax"] == v["btmax"] and base[0]["index"] < v["index"]
    path = tmp_path / "t.tsv"
    path.write_text(f"j2d5pt 0 0 1000 1000 0 {v['index']} {v['btmax']} 64 1.0\n"
                    f"j2d5pt 0 0 2000 2000 0 {v['index']} 1 64 1.0\n")
    pyan5d.tune_clear()
    pyan5d.tune_load(str(path))
    out = str(tmp_path / "u.tsv")
    pyan5d.tune_save(out)
    caps = [int(l.split()[-1]) for l in open(out) if not l.startswith("#")]
    assert caps == [v["regcap"], 0]
    pyan5d.tune_clear()


@pytest.mark.gpu
def test_autotune_picks_a_measured_config_and_results_stay_exact(torch_cuda):
    import oracle_lib as O
    print("case autotune j2d5pt 600x2000", flush=True)
    pyan5d.tune_clear()
    cfg, gs = pyan5d.autotune("j2d5pt", [600, 2000], max_measured=8)
    assert gs > 0 and cfg["bt"] >= 1
    rng = np.random.default_rng(0)
    u = rng.random((602, 2002)).astype(np.float32)
    A = np.stack([u, u])
    B = A.copy()
    pyan5d.forward_host("j2d5pt", A, [600, 2000], 21)  # uses the tuned entry
    O.run("j2d5pt", B, [600, 2000], 21)
    assert np.array_equal(A[1], B[1])
    pyan5d.tune_clear()


def test_resolve_config_follows_flags_then_table_then_defaults(tmp_path):
    pyan5d.tune_clear()
    # defaults: first variant, per-stencil bt, sl sized by the library
    d = pyan5d.resolve_config("j2d5pt", [1000, 1000])
    assert d["variant"] == 0 and d["bt"] == 4 and d["sl"] >= 1 and d["mode"] == pyan5d.EXACT
    # explicit values win
    e = pyan5d.resolve_config("j2d5pt", [1000, 1000], bt=3, variant=0, sl=48)
    assert (e["variant"], e["bt"], e["sl"]) == (0, 3, 48)
    # a tuned entry is used for sizes within 8x; its sl only for the exact size
    (tmp_path / "t.tsv").write_text("j2d5pt 0 0 1000 1000 0 0 6 64 1234.5\n")
    pyan5d.tune_load(str(tmp_path / "t.tsv"))
    t = pyan5d.resolve_config("j2d5pt", [1000, 1000])
    assert (t["variant"], t["bt"], t["sl"]) == (0, 6, 64)
    near = pyan5d.resolve_config("j2d5pt", [1200, 1000])
    assert near["bt"] == 6
    assert near["sl"] == pyan5d.resolve_config("j2d5pt", [1200, 1000], bt=6, variant=0)["sl"]
    far = pyan5d.resolve_config("j2d5pt", [8000, 8000])
    assert far["bt"] == 4
    # fast mode has its own rows
    assert pyan5d.resolve_config("j2d5pt", [1000, 1000], mode=pyan5d.FAST)["bt"] == 4
    pyan5d.tune_clear()
    with pytest.raises(pyan5d.An5dError):
        pyan5d.resolve_config("j2d5pt", [1000, 1000], variant=99)


def test_tune_workload_sets_cover_every_stencil():
    from pyan5d import tune
    names = {w[0] for w in tune.workloads("all")}
    assert names == set(tune.STENCILS_2D + tune.STENCILS_3D)
    assert len(names) == pyan5d.lib().an5d_stencil_count()
    bench = tune.workloads("bench")
    assert ("star3d1r", np.float32, [512, 512, 512], (pyan5d.EXACT, pyan5d.FAST), 1000) in bench


@pytest.mark.gpu
def test_autotune_leaves_the_callers_grid_as_it_was(torch_cuda):
    torch = torch_cuda
    print("case autotune caller grid j2d5pt 300x500", flush=True)
    n = [300, 500]
    host = pyan5d.init_grid("j2d5pt", n, np.float32, mode=1)
    t = torch.from_numpy(host).cuda()
    before = t.clone()
    pyan5d.tune_clear()
    cfg, gs = pyan5d.autotune("j2d5pt", n, max_measured=4, tensor=t, timesteps=40)
    torch.cuda.synchronize()
    assert gs > 0
    assert torch.equal(t.view(torch.uint8), before.view(torch.uint8))
    pyan5d.tune_clear()
