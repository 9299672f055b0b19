"""Batches of independent problems (sfmm_batch_*): one batch launch (a replay of
the captured HIP graph) must give each problem exactly what its own eager
evaluation gives, on every launch; problems of the 256-problem batch workload go
through one launch and are checked against the oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from simplefmm_amd import api, workloads as W
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

PROBLEMS = [
    # generator, n, seed, p, cap, kind, ncharge, precision, m2l_kernel
    ("uniform_cube", 20000, 61, 6, 32, api.POTGRAD, 1, api.FP64, 0),
    ("clusters", 50000, 62, 8, 64, api.POTGRAD, 1, api.FP64, 0),
    ("plummer", 30000, 63, 10, 64, api.POTGRAD, 1, api.MIXED, 0),
    ("sphere_shell", 10000, 64, 4, 16, api.CURL, 3, api.FP64, 0),
    ("uniform_cube", 15000, 65, 12, 64, api.POTENTIAL, 2, api.FP64, 0),
    ("plummer", 5000, 66, 8, 24, api.POTGRAD, 1, api.FP64, 1),
]


def _nout(kind, ncharge):
    return {api.POTGRAD: 4, api.CURL: 3}.get(kind, ncharge)


@pytest.fixture(scope="module")
def problems():
    import torch
    dev = torch.device("cuda:0")
    out = []
    for gen, n, seed, p, cap, kind, nq, prec, mk in PROBLEMS:
        pos, q = getattr(W, gen)(n, seed=seed)
        qdim = 3 if kind == api.CURL else (nq if kind == api.POTENTIAL else 1)
        Q = np.random.default_rng(seed).uniform(-1, 1, (n, qdim)) if qdim > 1 else q
        part = api.OctreePartitioner(pos, api.Box(pos)).refine(20, cap)
        fmm = api.FMM3D(part, p, ncharge=max(nq, 1), precision=prec, m2l_kernel=mk)
        dQ = torch.from_numpy(np.ascontiguousarray(Q, dtype=np.float64)).to(dev)
        dU = torch.zeros((n, _nout(kind, nq)), dtype=torch.float64, device=dev)
        # eager result
        fmm.evaluate_device(kind, nq, dQ.data_ptr(), dU.data_ptr())
        fmm.synchronize()
        ref = dU.cpu().numpy().copy()
        out.append(dict(fmm=fmm, kind=kind, nq=nq, dQ=dQ, dU=dU, ref=ref))
    return out


@pytest.mark.parametrize("streams", [0, 3], ids=["tuned", "three"])
def test_batch_replays_match_eager(problems, streams):
    import torch
    print(f"batch streams={streams}", flush=True)
    b = api.Batch([(pb["fmm"], pb["kind"], pb["nq"], pb["dQ"].data_ptr(), pb["dU"].data_ptr()) for pb in problems],
                  streams=streams)
    assert 1 <= b.streams <= len(problems)
    # tuning may have changed the plans' near-field tile shape (which changes the
    # rounding, not the method): the eager reference is taken after the batch exists
    for pb in problems:
        pb["dU"].zero_()
        torch.cuda.synchronize()
        pb["fmm"].evaluate_device(pb["kind"], pb["nq"], pb["dQ"].data_ptr(), pb["dU"].data_ptr())
        pb["fmm"].synchronize()
        pb["ref"] = pb["dU"].cpu().numpy().copy()
    for rep in range(3):
        for pb in problems:
            pb["dU"].zero_()
        torch.cuda.synchronize()
        b.launch()
        b.synchronize()
        for k, pb in enumerate(problems):
            got = pb["dU"].cpu().numpy()
            assert np.array_equal(got, pb["ref"]), (rep, k)
        # eager work between replays: an ordinary evaluation of one of the plans
        pb = problems[rep % len(problems)]
        scratch = torch.zeros_like(pb["dU"])
        pb["fmm"].evaluate_device(pb["kind"], pb["nq"], pb["dQ"].data_ptr(), scratch.data_ptr())
        pb["fmm"].synchronize()
        assert np.array_equal(scratch.cpu().numpy(), pb["ref"])
    del b


def test_batch_rejects_a_repeated_plan(problems):
    pb = problems[0]
    item = (pb["fmm"], pb["kind"], pb["nq"], pb["dQ"].data_ptr(), pb["dU"].data_ptr())
    with pytest.raises(api.SfmmError, match="twice"):
        api.Batch([item, item], streams=1)


def test_batch256_problems_against_oracle():
    # one problem of each order from the 256-problem batch workload (clusters, leaf
    # capacity 64), the smallest of its order, all in one tuned batch launch: FP64
    # against the oracle at 1e-12; the order-12 problem in the mixed path against
    # FP64 and the FP64 direct sum: its difference from FP64 is within the
    # truncation error of order p (FP64's own error against the direct sum), or
    # the 22-bit rounding bound where that is smaller
    import torch
    dev = torch.device("cuda:0")
    specs = W.batch_specs(256)
    chosen = {}
    for n, p, seed in specs:
        if p not in chosen or n < chosen[p][0]:
            chosen[p] = (n, p, seed)
    assert sorted(chosen) == [6, 8, 10, 12]
    probs = []
    for n, p, seed in sorted(chosen.values()):
        pos, q = W.clusters(n, seed=seed)
        part = api.OctreePartitioner(pos, api.Box(pos)).refine(20, 64)
        probs.append(dict(n=n, p=p, pos=pos, q=q, part=part, fmm=api.FMM3D(part, p, ncharge=1),
                          dq=torch.from_numpy(q).to(dev), du=torch.zeros((n, 4), dtype=torch.float64, device=dev)))
    print("batch256 subset: " + ", ".join(f"n={pb['n']} p={pb['p']}" for pb in probs), flush=True)
    b = api.Batch([(pb["fmm"], api.POTGRAD, 1, pb["dq"].data_ptr(), pb["du"].data_ptr()) for pb in probs],
                  streams=0)
    b.launch()
    b.synchronize()
    for pb in probs:
        ref = O.Tree(pb["pos"], 20, 64).fmm(pb["p"], pb["q"], O.POTGRAD)
        got = pb["du"].cpu().numpy()
        err = rel_l2(got, ref)
        print(f"  n={pb['n']} p={pb['p']}: rel L2 vs oracle {err:.2e}", flush=True)
        assert err < 1e-12
    big = next(pb for pb in probs if pb["p"] == 12)
    mixed = api.FMM3D(big["part"], big["p"], ncharge=1, precision=api.MIXED)
    du = torch.zeros((big["n"], 4), dtype=torch.float64, device=dev)
    bm = api.Batch([(mixed, api.POTGRAD, 1, big["dq"].data_ptr(), du.data_ptr())], streams=0)
    bm.launch()
    bm.synchronize()
    U = du.cpu().numpy()
    V = big["du"].cpu().numpy()
    tg = np.sort(np.random.default_rng(3).choice(big["n"], 2000, replace=False))
    ref = mixed.direct(api.POTGRAD, big["q"], tg)
    for name, cols in (("phi", slice(0, 1)), ("grad", slice(1, 4))):
        e_mix = rel_l2(U[tg, cols], ref[:, cols])
        e_fp64 = rel_l2(V[tg, cols], ref[:, cols])
        diff = rel_l2(U[:, cols], V[:, cols])
        bound = max(e_fp64, 5e-7)
        print(f"  mixed n={big['n']} p={big['p']} {name}: {e_mix:.2e} against the direct sum (FP64 {e_fp64:.2e}),"
              f" {diff:.2e} against FP64", flush=True)
        assert diff <= bound and e_mix <= e_fp64 + bound, name
