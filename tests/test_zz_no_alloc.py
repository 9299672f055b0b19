"""sfmm_direct_device_noalloc (static scratch, no device allocation) against
sfmm_direct on the same plan: bit for bit, on two calls in a row, the second
with fewer targets, for each kind of charge."""
import numpy as np
import pytest

from simplefmm_amd import api, workloads as W

pytestmark = pytest.mark.gpu

CASES = [
    # kind, ncharge, ntarget of the first and second call
    (api.POTGRAD, 1, 5000, 700),
    (api.POTENTIAL, 2, 3000, 64),
    (api.CURL, 3, 2000, 1),
]


@pytest.fixture(scope="module")
def plan():
    pos, q = W.plummer(40000, seed=71)
    part = api.OctreePartitioner(pos, api.Box(pos)).refine(20, 64)
    return pos, q, api.FMM3D(part, 6, ncharge=3)


@pytest.mark.parametrize("kind,nq,n1,n2", CASES, ids=["potgrad", "pot2", "curl"])
def test_noalloc_direct_matches_direct(plan, kind, nq, n1, n2):
    import torch
    pos, q, fmm = plan
    n = len(pos)
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(kind * 10 + nq)
    qdim = 3 if kind == api.CURL else (nq if kind == api.POTENTIAL else 1)
    Q = q if qdim == 1 else rng.uniform(-1, 1, (n, qdim))
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    nout = {api.POTGRAD: 4, api.CURL: 3}.get(kind, nq)
    dQ = torch.from_numpy(Q).to(dev)
    for call, nt in enumerate((n1, n2)):
        targets = np.sort(rng.choice(n, nt, replace=False)).astype(np.int64)
        want = fmm.direct(kind, Q, targets, ncharge=nq)
        dT = torch.from_numpy(targets).to(dev)
        dO = torch.full((nt, nout), 7.0, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        print(f"noalloc direct kind={kind} nq={nq} call={call} ntarget={nt}", flush=True)
        fmm.direct_device_noalloc(kind, nq, dQ.data_ptr(), nt, dT.data_ptr(), dO.data_ptr())
        fmm.synchronize()
        got = dO.cpu().numpy()
        assert np.array_equal(got, want), f"call {call}: max diff {np.max(np.abs(got - want))}"


def test_noalloc_direct_rejects_too_many_targets(plan):
    import torch
    _, q, fmm = plan
    dQ = torch.from_numpy(np.ascontiguousarray(q, dtype=np.float64)).to("cuda:0")
    with pytest.raises(api.SfmmError, match="static scratch"):
        fmm.direct_device_noalloc(api.POTENTIAL, 1, dQ.data_ptr(), api.DIRECT_NOALLOC_MAX_TARGETS + 1, 0, 0)


def test_noalloc_direct_bad_target_gives_nan(plan):
    import torch
    _, q, fmm = plan
    dev = torch.device("cuda:0")
    dQ = torch.from_numpy(np.ascontiguousarray(q, dtype=np.float64)).to(dev)
    dT = torch.tensor([0, len(q), -1, 5], dtype=torch.int64, device=dev)
    dO = torch.zeros((4, 1), dtype=torch.float64, device=dev)
    fmm.direct_device_noalloc(api.POTENTIAL, 1, dQ.data_ptr(), 4, dT.data_ptr(), dO.data_ptr())
    fmm.synchronize()
    got = dO.cpu().numpy()[:, 0]
    assert np.isfinite(got[[0, 3]]).all() and np.isnan(got[[1, 2]]).all()
    want = fmm.direct(api.POTENTIAL, q, np.array([0, 5], np.int64))[:, 0]
    assert np.array_equal(got[[0, 3]], want)
