"""Batches of independent problems (sfmm_batch_*): one batch launch must give each
problem exactly what its own eager evaluation gives, on every launch."""
import numpy as np
import pytest

from simplefmm_amd import api, workloads as W

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
    assert 1 <= b.streams <= 8
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
