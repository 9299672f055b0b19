"""The boundary on the GPU: the C++ drop-in and its driver against the reference's
fixtures and the oracle, device trees against the reference's list digests,
ordering between a plan's own stream and batch launches, recovery from failed
device allocations, and the mixed path over every kind and a spread of orders."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from simplefmm_amd import api, workloads as W
from tests.helpers import golden_cases, list_sha, load_golden, rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "simplefmm_amd", "cpp", "sfmm_driver")
LIST_KEYS = ["c_node", "leaf2node", "idx_row", "idx_val", "ilist_row", "ilist_val", "clist_row", "clist_val",
             "plist", "nlist_row", "nlist_val"]


def _read_driver_output(path):
    with open(path, "rb") as f:
        assert f.read(4) == b"SFMR"
        f.read(4)
        n = int(np.frombuffer(f.read(8), np.int64)[0])
        nout = int(np.frombuffer(f.read(4), np.int32)[0])
        f.read(8)
        return np.frombuffer(f.read(), np.float64).reshape(n, nout)


def test_driver_matches_reference_fixture(tmp_path):
    # sfmm_driver runs FMM3D<double, 3, 8, Vector, true>: the shape of the
    # vector_5k_p8 fixture (main.cpp's rinv and rinv<true>)
    assert os.path.exists(DRIVER), "build the driver: make -C simplefmm_amd/cpp"
    meta, g, pos, Q = load_golden("vector_5k_p8")
    assert meta["p"] == 8
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    W.write_harness_input(str(fin), pos, Q)
    print("driver on vector_5k_p8", flush=True)
    r = subprocess.run([DRIVER, "0", str(meta["cap"]), "200", "1", "--in", str(fin), "--out", str(fout)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = _read_driver_output(fout)
    idx = g["idx"]
    assert rel_l2(out[idx, :3], g["pot_full"]) < 1e-12
    assert rel_l2(out[idx, 3:], g["curl_full"]) < 1e-12


@pytest.mark.parametrize("case", golden_cases(None, self_lists=None))
def test_device_tree_matches_reference_digests(case):
    meta, _, pos, _ = load_golden(case)
    print(f"device tree {case} n={meta['n']}", flush=True)
    part = api.OctreePartitioner(pos, api.Box(pos)).refine(meta["max_level"], meta["cap"], on_device=True)
    assert part.level() == meta["depth"]
    for l in range(meta["depth"] + 1):
        ref = meta["levels"][str(l)]
        mine = part.level_lists(l)
        assert mine["n_node"] == ref["n_node"] and mine["n_leaf"] == ref["n_leaf"]
        for k in LIST_KEYS:
            assert list_sha(mine[k]) == ref[k], f"level {l} list {k}"


def _self_lists(pos, seed):
    from tests.helpers import golden_self_lists
    return golden_self_lists({"seed": seed}, pos)


def test_cpp_dropin_interface(tmp_path):
    # comm(), slist(), self_list() and a plain lambda as is_self, through the header
    exe = tmp_path / "dropin_api"
    cmd = ["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "simplefmm_amd", "cpp"),
           os.path.join(ROOT, "tests", "dropin_api.cpp"), "-o", str(exe), "-L", os.path.join(ROOT, "simplefmm_amd"),
           "-lsfmm", "-Wl,-rpath," + os.path.join(ROOT, "simplefmm_amd")]
    subprocess.run(cmd, check=True, timeout=300)
    n = 4000
    pos, _ = W.clusters(n, seed=81)
    Q = np.random.default_rng(81).uniform(-1, 1, (n, 3))
    row, val = _self_lists(pos, 81)
    fin, fsl, fout = tmp_path / "in.bin", tmp_path / "self.bin", tmp_path / "out.bin"
    W.write_harness_input(str(fin), pos, Q)
    with open(fsl, "wb") as f:
        f.write(b"SFMS" + np.int64(n).tobytes() + np.int64(len(val)).tobytes() + row.tobytes() + val.tobytes())
    print("dropin_api run", flush=True)
    r = subprocess.run([str(exe), str(fin), str(fsl), str(fout)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    with open(fout, "rb") as f:
        assert f.read(4) == b"SFMD"
        assert int(np.frombuffer(f.read(8), np.int64)[0]) == n
        outs = [np.frombuffer(f.read(24 * n), np.float64).reshape(n, 3) for _ in range(6)]
        nnz = int(np.frombuffer(f.read(8), np.int64)[0])
        srow = np.frombuffer(f.read(8 * (n + 1)), np.int64)
        sval = np.frombuffer(f.read(8 * nnz), np.int64)
    pot_dev, pot_host, curl_dev, curl_host, near_host, far = outs
    t = O.Tree(pos, 20, 24)
    lists = (row, val)
    assert rel_l2(pot_dev, t.fmm(6, Q, O.VECPOT, O.BOTH, self_lists=lists)) < 1e-12
    assert rel_l2(curl_dev, t.fmm(6, Q, O.CURL, O.BOTH, self_lists=lists)) < 1e-12
    assert rel_l2(near_host, t.fmm(6, Q, O.CURL, O.NEAR, self_lists=lists)) < 1e-12
    assert rel_l2(far, t.fmm(6, Q, O.VECPOT, O.FAR)) < 1e-12
    # the lambda is evaluated on the host over the near pairs: the same sums
    assert rel_l2(pot_host, pot_dev) < 1e-13 and rel_l2(curl_host, curl_dev) < 1e-13
    # slist: the listed entries outside the near field, sorted, duplicates kept
    exp_row, exp_val = [0], []
    for i in range(n):
        exp_val.extend(sorted(int(j) for j in val[row[i]:row[i + 1]] if not t.is_neighbor(i, j)))
        exp_row.append(len(exp_val))
    assert nnz > 100 and np.array_equal(srow, exp_row) and np.array_equal(sval, exp_val)


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    return torch, torch.device("cuda:0")


@pytest.mark.parametrize("streams", [1, 3])
def test_batch_is_ordered_with_the_plans_own_stream(torch_dev, streams):
    # eager evaluations and batch launches on the same plans, with no
    # synchronisation between them, must not overwrite each other's buffers
    torch, dev = torch_dev
    probs = []
    for k, (gen, n, p) in enumerate([("plummer", 40000, 8), ("clusters", 30000, 10), ("uniform_cube", 20000, 6)]):
        pos, q = getattr(W, gen)(n, seed=90 + k)
        part = api.OctreePartitioner(pos, api.Box(pos)).refine(20, 48)
        fmm = api.FMM3D(part, p, ncharge=1)
        dq = torch.from_numpy(q).to(dev)
        ref = np.zeros((n, 4))
        fmm.rinv_potgrad(q, ref)
        probs.append(dict(fmm=fmm, dq=dq, ref=ref, a=torch.zeros((n, 4), dtype=torch.float64, device=dev),
                          b=torch.zeros((n, 4), dtype=torch.float64, device=dev)))
    batch = api.Batch([(pb["fmm"], api.POTGRAD, 1, pb["dq"].data_ptr(), pb["b"].data_ptr()) for pb in probs],
                      streams=streams)
    for rep in range(3):
        for pb in probs:
            pb["a"].zero_()
            pb["b"].zero_()
        torch.cuda.synchronize()
        for pb in probs:  # eager, then the batch, then eager again, all asynchronous
            pb["fmm"].evaluate_device(api.POTGRAD, 1, pb["dq"].data_ptr(), pb["a"].data_ptr())
        batch.launch()
        for pb in probs:
            pb["fmm"].evaluate_device(api.POTGRAD, 1, pb["dq"].data_ptr(), pb["a"].data_ptr())
        batch.synchronize()
        for pb in probs:
            pb["fmm"].synchronize()
        for k, pb in enumerate(probs):
            assert np.array_equal(pb["b"].cpu().numpy(), pb["ref"]), (rep, k)
            assert np.array_equal(pb["a"].cpu().numpy(), 2 * pb["ref"]), (rep, k)


def test_failed_allocations_are_reported_and_recoverable(torch_dev, monkeypatch):
    # SFMM_FAULT_ALLOC=k lets k device allocations succeed and fails the rest:
    # every failure point of plan creation and of an evaluation's first use must
    # return SFMM_ERR_ALLOC (twice in a row, never launching on partial state)
    # and the plan must work once allocations succeed again
    torch, dev = torch_dev
    pos, q = W.clusters(6000, seed=95)
    part = api.OctreePartitioner(pos, api.Box(pos)).refine(20, 32)
    ref = np.zeros((6000, 4))
    api.FMM3D(part, 7, ncharge=1).rinv_potgrad(q, ref)
    dq = torch.from_numpy(q).to(dev)
    outcomes = set()
    for k in range(0, 90, 3):
        print(f"fault after {k} allocations", flush=True)
        monkeypatch.setenv("SFMM_FAULT_ALLOC", str(k))
        try:
            fmm = api.FMM3D(part, 7, ncharge=1)
        except api.SfmmError as e:
            assert "sfmm error 3" in str(e)
            outcomes.add("create")
            continue
        finally:
            monkeypatch.delenv("SFMM_FAULT_ALLOC", raising=False)
        monkeypatch.setenv("SFMM_FAULT_ALLOC", "0")
        du = torch.zeros((6000, 4), dtype=torch.float64, device=dev)
        for attempt in range(2):
            with pytest.raises(api.SfmmError, match="sfmm error 3"):
                fmm.evaluate_device(api.POTGRAD, 1, dq.data_ptr(), du.data_ptr())
        outcomes.add("evaluate")
        monkeypatch.delenv("SFMM_FAULT_ALLOC")
        fmm.evaluate_device(api.POTGRAD, 1, dq.data_ptr(), du.data_ptr())
        fmm.synchronize()
        assert np.array_equal(du.cpu().numpy(), ref)
    assert outcomes == {"create", "evaluate"}


def test_failed_allocations_midway_through_the_work_lists(torch_dev, monkeypatch):
    # failures between the uploads of the M2L / M2M / L2L work lists
    torch, dev = torch_dev
    pos, q = W.plummer(5000, seed=96)
    part = api.OctreePartitioner(pos, api.Box(pos)).refine(20, 24)
    ref = np.zeros((5000, 4))
    api.FMM3D(part, 6, ncharge=1).rinv_potgrad(q, ref)
    dq = torch.from_numpy(q).to(dev)
    for k in range(0, 24):
        fmm = api.FMM3D(part, 6, ncharge=1)
        du = torch.zeros((5000, 4), dtype=torch.float64, device=dev)
        monkeypatch.setenv("SFMM_FAULT_ALLOC", str(k))
        failed = 0
        for attempt in range(2):
            try:
                fmm.evaluate_device(api.POTGRAD, 1, dq.data_ptr(), du.data_ptr())
            except api.SfmmError as e:
                assert "sfmm error 3" in str(e)
                failed += 1
        monkeypatch.delenv("SFMM_FAULT_ALLOC")
        fmm.synchronize()
        du.zero_()
        fmm.evaluate_device(api.POTGRAD, 1, dq.data_ptr(), du.data_ptr())
        fmm.synchronize()
        assert np.array_equal(du.cpu().numpy(), ref), k
        if failed == 0:
            break
    assert k > 3


@pytest.mark.parametrize("p", [4, 6, 10, 16])
def test_mixed_every_kind_against_oracle(p):
    # every charge count (1..4 columns per target in the M2L blocks) and the curl;
    # p <= 12 runs the split f16 M2L, p = 16 the FP32 stages
    n = 5000
    pos, _ = W.clusters(n, seed=100 + p)
    part = api.OctreePartitioner(pos, api.Box(pos)).refine(20, 32)
    fmm = api.FMM3D(part, p, ncharge=4, precision=api.MIXED)
    t = O.Tree(pos, 20, 32)
    rng = np.random.default_rng(p)
    # the FP64 oracle's own truncation error at this order, for scale
    print(f"mixed p={p}", flush=True)
    for nc in (1, 2, 3, 4):
        Q = rng.uniform(-1, 1, (n, nc))
        U = np.zeros((n, nc))
        fmm._eval(api.POTENTIAL, nc, Q, U, api.BOTH)
        ref = np.stack([t.fmm(p, Q[:, c].copy(), O.POT) for c in range(nc)], axis=1)
        err = rel_l2(U, ref)
        print(f"  potential nc={nc}: {err:.3e}", flush=True)
        assert err < 1e-6, nc
    Q3 = rng.uniform(-1, 1, (n, 3))
    F = np.zeros((n, 3))
    fmm.rinv(Q3, F, gradient=True)
    err = rel_l2(F, t.fmm(p, Q3, O.CURL))
    print(f"  curl: {err:.3e}", flush=True)
    assert err < 1e-6
    q = rng.uniform(-1, 1, n)
    U = np.zeros((n, 4))
    fmm.rinv_potgrad(q, U)
    err = rel_l2(U, t.fmm(p, q, O.POTGRAD))
    print(f"  potgrad: {err:.3e}", flush=True)
    assert err < 1e-6


@pytest.mark.parametrize("precision", [api.FP64, api.MIXED], ids=["fp64", "mixed"])
def test_tuned_tiles_keep_the_results(torch_dev, precision):
    # sfmm_plan_tune picks the tile shapes by timing; whatever it picks, the
    # evaluation stays the method: FP64 against the oracle at 1e-12, mixed within
    # its rounding of FP64
    torch, dev = torch_dev
    pos, q = W.plummer(30000, seed=97)
    part = api.OctreePartitioner(pos, api.Box(pos)).refine(20, 64)
    fmm = api.FMM3D(part, 10, ncharge=1, precision=precision)
    dq = torch.from_numpy(q).to(dev)
    print(f"tune precision={precision}", flush=True)
    fmm.tune(api.POTGRAD, 1, dq.data_ptr())
    tiles = fmm.tiles()
    assert tiles["p2p_targets_per_lane"] in (1, 2, 4)
    assert tiles["m2l_waves"] in ((4, 8) if precision == api.MIXED else (0,))
    U = np.zeros((30000, 4))
    fmm.rinv_potgrad(q, U)
    ref = O.Tree(pos, 20, 64).fmm(10, q, O.POTGRAD)
    err = rel_l2(U, ref)
    print(f"  tiles {tiles}: rel L2 vs oracle {err:.2e}", flush=True)
    assert err < (1e-12 if precision == api.FP64 else 1e-6)


@pytest.mark.parametrize("tpl,waves", [(1, 8), (4, 8), (2, 4), (1, 4)])
def test_every_tile_shape_against_oracle(torch_dev, tpl, waves):
    # every near-field and dense-M2L tile shape the tuner can pick: FP64 against the
    # oracle at 1e-12, mixed within its rounding, for potential+gradient and a
    # three-component potential
    torch, dev = torch_dev
    pos, q = W.clusters(12000, seed=98)
    part = api.OctreePartitioner(pos, api.Box(pos)).refine(20, 40)
    tree = O.Tree(pos, 20, 40)
    Q3 = np.random.default_rng(5).uniform(-1, 1, (12000, 3))
    ref_pg = tree.fmm(8, q, O.POTGRAD)
    ref_p3 = np.stack([tree.fmm(8, Q3[:, c].copy(), O.POT) for c in range(3)], axis=1)
    for prec in (api.FP64, api.MIXED):
        fmm = api.FMM3D(part, 8, ncharge=3, precision=prec)
        fmm.set_tiles(tpl, waves)
        assert fmm.tiles()["p2p_targets_per_lane"] == tpl
        print(f"tiles tpl={tpl} waves={waves} precision={prec}", flush=True)
        U = np.zeros((12000, 4))
        fmm.rinv_potgrad(q, U)
        V = np.zeros((12000, 3))
        fmm._eval(api.POTENTIAL, 3, Q3, V, api.BOTH)
        tol = 1e-12 if prec == api.FP64 else 2e-6
        assert rel_l2(U, ref_pg) < tol
        assert rel_l2(V, ref_p3) < tol
