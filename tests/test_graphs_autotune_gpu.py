"""HIP-graph capture of plan executions and the permanent autotuner."""
import numpy as np
import pytest

import oracle
from perceval_amd.utils.workloads import submatrix_batch

pytestmark = pytest.mark.gpu


def test_graph_replays_match_eager(gpu):
    import torch
    from perceval_amd.graphs import graphed_permanents
    from perceval_amd.permanent import PermanentPlan, pack
    _, orders, mats = submatrix_batch(256, 9, 18, seed=3)
    o, flat = pack(mats)
    plan = PermanentPlan(o)
    d = torch.from_numpy(flat).cuda()
    eager = plan(d).clone()
    out = torch.zeros_like(eager)
    g = graphed_permanents(plan, d, out)
    for rep in range(4):
        out.zero_()
        _ = torch.randn(1000, device="cuda").sum().item()   # unrelated eager work between replays
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, eager), f"replay {rep} differs"
    ref = oracle.perm_batch(mats[:8], precise=True)
    assert np.max(np.abs(eager[:8].cpu().numpy() - ref) / np.abs(ref)) < 1e-10
    plan.close()


def test_graph_slos_replays(gpu):
    import torch
    from perceval_amd.graphs import graphed_slos
    from perceval_amd.slos import COEFS, PROBS, SlosPlan
    from perceval_amd.utils.workloads import haar_unitary
    for inp, mode in (([1, 1, 1, 0, 0, 0, 0], PROBS), ([1, 0, 2, 0, 1, 1, 0, 0, 1], COEFS)):
        plan = SlosPlan(inp)
        dU = torch.from_numpy(haar_unitary(len(inp), seed=len(inp))).cuda()
        eager = plan.run(dU, mode).clone()
        out = torch.empty_like(eager)
        g = graphed_slos(plan, dU, mode, out, repeats=2)
        for rep in range(3):
            out.fill_(-1)
            _ = torch.randn(1000, device="cuda").sum().item()   # unrelated eager work between replays
            g.replay()
            torch.cuda.synchronize()
            assert torch.equal(out, eager), f"replay {rep} differs"
        g.close()
        plan.close()


def test_graph_rejects_bad_arguments(gpu):
    import ctypes
    import torch
    from perceval_amd import _native
    from perceval_amd.permanent import PermanentPlan
    lib = _native.load()
    plan = PermanentPlan([5, 6])
    d = torch.zeros(61, dtype=torch.complex128, device="cuda")
    out = torch.zeros(2, dtype=torch.complex128, device="cuda")
    g = ctypes.c_void_p()
    assert lib.pcvl_perm_graph_create(plan._plan, _native.ptr(d), _native.ptr(out), 0, ctypes.byref(g)) == -1
    assert lib.pcvl_perm_graph_create(None, _native.ptr(d), _native.ptr(out), 1, ctypes.byref(g)) == -1
    assert lib.pcvl_graph_launch(None, None) == -1
    assert lib.pcvl_graph_destroy(None) == 0
    plan.close()


def test_autotune_picks_a_measured_config(gpu, tmp_path, monkeypatch):
    import perceval_amd.autotune as at
    monkeypatch.setattr(at, "_CACHE_PATH", str(tmp_path / "tune.json"))
    res = at.tune_permanent(16, batch=512, reps=2)
    assert res["variant"] in at.VARIANTS and res["chunk_log2"] <= 9
    assert res["ms"] == min(t["ms"] for t in res["trials"])
    again = at.tune_permanent(16, batch=512)
    assert again["variant"] == res["variant"] and again["chunk_log2"] == res["chunk_log2"]
    plan = at.tuned_plan([16] * 512)
    assert plan.info()["single_order"] == 16
    plan.close()


def test_autotune_mixed_and_slos(gpu, tmp_path, monkeypatch):
    import numpy as np
    import torch
    import oracle
    import perceval_amd.autotune as at
    from perceval_amd.permanent import pack
    from perceval_amd.slos import PROBS
    from perceval_amd.utils.workloads import haar_unitary
    monkeypatch.setattr(at, "_CACHE_PATH", str(tmp_path / "tune.json"))
    orders = np.array([9, 12, 14, 10, 13, 11] * 20, dtype=np.int32)
    res = at.tune_permanent_mixed(orders, reps=2)
    assert res["chunk_log2"] in at.CHUNKS and res["ms"] == min(t["ms"] for t in res["trials"])
    rng = np.random.RandomState(3)
    mats = [rng.randn(k, k) + 1j * rng.randn(k, k) for k in orders[:12]]
    o, flat = pack(mats)
    plan = at.tuned_plan(o)
    got = plan(torch.from_numpy(flat).cuda()).cpu().numpy()
    plan.close()
    ref = oracle.perm_batch(mats, precise=True)
    assert np.max(np.abs(got - ref) / np.abs(ref)) < 1e-10
    inp = [1, 1, 1, 0, 0, 1, 0, 0, 0, 0]
    t = at.tune_slos(inp, reps=2)
    assert t["kernel"] in at.SLOS_KERNELS and t["ms"] == min(x["ms"] for x in t["trials"])
    assert at.tune_slos(inp)["kernel"] == t["kernel"]          # cached
    sp = at.tuned_slos_plan(inp)
    U = haar_unitary(10, seed=8)
    p = sp.run(U, PROBS).cpu().numpy()
    sp.close()
    assert np.max(np.abs(p - oracle.slos_probs(U, inp))) < 1e-12
