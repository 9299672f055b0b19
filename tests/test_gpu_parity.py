"""HIP path (libsfmm.so on the GPU) against the reference outputs, the oracle and
the FP64 direct sum.  Every comparison goes through the C-ABI.

Bars: 1e-12 relative L2 against the reference FMM (golden fixtures) and the
oracle on the same inputs; direct-sum agreement within the scheme's truncation
error at full size; bitwise repeatability; exact accumulation semantics.
"""
import numpy as np
import pytest

from oracle import oracle as O
from simplefmm_amd import api, workloads as W
from tests.helpers import golden_cases, load_golden, rel_l2

pytestmark = pytest.mark.gpu

PARITY = 1e-12


def _plan(pos, cap, p, ncharge=1, max_level=20):
    part = api.OctreePartitioner(pos, api.Box(pos)).refine(max_level, cap)
    return part, api.FMM3D(part, p, ncharge)


@pytest.mark.parametrize("case", golden_cases("potgrad"))
def test_potgrad_matches_reference(case):
    meta, g, pos, q = load_golden(case)
    print(f"case {case} n={meta['n']} p={meta['p']}", flush=True)
    _, fmm = _plan(pos, meta["cap"], meta["p"], 1, meta["max_level"])
    idx = g["idx"]
    for which, suffix in ((api.BOTH, "full"), (api.FAR, "far"), (api.NEAR, "near")):
        if f"phi_{suffix}" not in g:
            continue
        U = np.zeros((meta["n"], 4))
        fmm.rinv_potgrad(q, U, which)
        assert rel_l2(U[idx, 0], g[f"phi_{suffix}"]) < PARITY, suffix
        assert rel_l2(U[idx, 1:], g[f"grad_{suffix}"]) < PARITY, suffix


def test_vector_and_curl_match_reference():
    meta, g, pos, Q = load_golden("vector_5k_p8")
    _, fmm = _plan(pos, meta["cap"], meta["p"], 3, meta["max_level"])
    idx = g["idx"]
    for which, suffix in ((api.BOTH, "full"), (api.FAR, "far"), (api.NEAR, "near")):
        U = np.zeros((meta["n"], 3))
        F = np.zeros((meta["n"], 3))
        fmm.rinv(Q, U, which=which)
        fmm.rinv(Q, F, gradient=True, which=which)
        assert rel_l2(U[idx], g[f"pot_{suffix}"]) < PARITY, suffix
        assert rel_l2(F[idx], g[f"curl_{suffix}"]) < PARITY, suffix


ORACLE_CASES = [
    ("uniform", 3000, 16, 5),
    ("clusters", 4000, 10, 7),
    ("plummer", 3000, 12, 9),
    ("sphere", 3000, 8, 3),
    ("uniform", 60, 64, 6),     # depth 0: near field only
    ("uniform", 400, 64, 6),    # depth 1: near field only
    ("uniform", 5000, 2, 4),    # max_level 4 below: fat leaves (> 64 particles)
    ("uniform", 2500, 32, 16),  # highest order
    ("uniform", 2500, 32, 1),   # lowest order
]


def _pos(kind, n, seed):
    return {"uniform": W.uniform_cube, "clusters": W.clusters, "plummer": W.plummer,
            "sphere": W.sphere_shell}[kind](n, seed=seed)[0]


@pytest.mark.parametrize("m2l_kernel", [0, 1], ids=["mfma", "rows"])
@pytest.mark.parametrize("kind,n,cap,p", ORACLE_CASES)
def test_all_kinds_match_oracle(kind, n, cap, p, m2l_kernel):
    print(f"case {kind} n={n} cap={cap} p={p} m2l_kernel={m2l_kernel}", flush=True)
    pos = _pos(kind, n, 41)
    max_level = 4 if cap == 2 else 20
    part = api.OctreePartitioner(pos, api.Box(pos)).refine(max_level, cap)
    fmm = api.FMM3D(part, p, ncharge=4, m2l_kernel=m2l_kernel)
    ot = O.Tree(pos, max_level, cap)
    rng = np.random.default_rng(n + p)
    q1 = rng.uniform(-1, 1, n)
    U = np.zeros((n, 4))
    fmm.rinv_potgrad(q1, U)
    assert rel_l2(U, ot.fmm(p, q1, O.POTGRAD)) < PARITY
    for nc in (1, 2, 4):
        Q = rng.uniform(-1, 1, (n, nc))
        U = np.zeros((n, nc))
        fmm._eval(api.POTENTIAL, nc, Q, U, api.BOTH)
        ref = np.zeros((n, nc))
        for c in range(nc):
            ref[:, c] = ot.fmm(p, Q[:, c].copy(), O.POT)
        assert rel_l2(U, ref) < PARITY, nc
    Q3 = rng.uniform(-1, 1, (n, 3))
    F = np.zeros((n, 3))
    fmm.rinv(Q3, F, gradient=True)
    assert rel_l2(F, ot.fmm(p, Q3, O.CURL)) < PARITY


def test_outputs_accumulate_into_U():
    pos, q = W.uniform_cube(3000, seed=42)
    _, fmm = _plan(pos, 24, 6)
    U0 = np.random.default_rng(1).normal(size=(3000, 4))
    U = U0.copy()
    fmm.rinv_potgrad(q, U)
    V = np.zeros((3000, 4))
    fmm.rinv_potgrad(q, V)
    assert np.max(np.abs(U - (U0 + V))) < 1e-12 * np.max(np.abs(V))


def test_repeat_is_bitwise_identical():
    pos, q = W.plummer(20000, seed=43)
    _, fmm = _plan(pos, 64, 8)
    a = np.zeros((20000, 4))
    b = np.zeros((20000, 4))
    fmm.rinv_potgrad(q, a)
    fmm.rinv_potgrad(q, b)
    assert np.array_equal(a, b)


def test_direct_kernel_matches_oracle():
    pos, q = W.clusters(20000, seed=44)
    _, fmm = _plan(pos, 64, 4)
    tg = np.arange(5, 20000, 37)
    assert rel_l2(fmm.direct(api.POTGRAD, q, tg), O.direct(pos, q, tg)) < 1e-14


def test_full_size_uniform_against_direct_sum():
    # 2^20 points, p=10: the FMM agrees with the FP64 direct sum to its truncation
    # error, and is linear in the charges to rounding
    n = 1 << 20
    pos, q = W.uniform_cube(n, seed=45)
    _, fmm = _plan(pos, 64, 10)
    U = np.zeros((n, 4))
    fmm.rinv_potgrad(q, U)
    tg = np.random.default_rng(2).choice(n, 2000, replace=False)
    ref = fmm.direct(api.POTGRAD, q, tg)
    assert rel_l2(U[tg, 0], ref[:, 0]) < 5e-6
    assert rel_l2(U[tg, 1:], ref[:, 1:]) < 3e-5
    q2 = np.random.default_rng(3).uniform(-1, 1, n)
    V = np.zeros((n, 4))
    fmm.rinv_potgrad(q2, V)
    Wc = np.zeros((n, 4))
    fmm.rinv_potgrad(2.0 * q - 0.5 * q2, Wc)
    assert rel_l2(Wc, 2.0 * U - 0.5 * V) < 1e-12


def test_mfma_and_row_kernels_agree_at_size():
    pos, q = W.plummer(200000, seed=46)
    part = api.OctreePartitioner(pos, api.Box(pos)).refine(20, 64)
    a = np.zeros((200000, 4))
    b = np.zeros((200000, 4))
    api.FMM3D(part, 12, ncharge=1, m2l_kernel=0).rinv_potgrad(q, a, api.FAR)
    api.FMM3D(part, 12, ncharge=1, m2l_kernel=1).rinv_potgrad(q, b, api.FAR)
    assert rel_l2(a, b) < 1e-13


@pytest.mark.parametrize("p", [8, 12])
def test_mixed_precision_close_to_fp64(p):
    pos, q = W.plummer(100000, seed=47)
    part = api.OctreePartitioner(pos, api.Box(pos)).refine(20, 64)
    a = np.zeros((100000, 4))
    b = np.zeros((100000, 4))
    api.FMM3D(part, p, ncharge=1).rinv_potgrad(q, a)
    api.FMM3D(part, p, ncharge=1, precision=api.MIXED).rinv_potgrad(q, b)
    err_phi, err_grad = rel_l2(b[:, 0], a[:, 0]), rel_l2(b[:, 1:], a[:, 1:])
    print(f"mixed vs fp64 p={p}: phi {err_phi:.3e} grad {err_grad:.3e}", flush=True)
    assert err_phi < 1e-6 and err_grad < 1e-6


def _self_lists(pos, rng):
    """self_generator(i) = {i, nearest neighbour of i, one random particle}."""
    n = len(pos)
    d2 = ((pos[:, None, :] - pos[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d2, np.inf)
    nn = np.argmin(d2, axis=1)
    rnd = rng.integers(0, n, n)
    row, val = [0], []
    for i in range(n):
        s = sorted({i, int(nn[i]), int(rnd[i])})
        val.extend(s)
        row.append(len(val))
    return np.array(row, np.int64), np.array(val, np.int64)


@pytest.mark.parametrize("m2l_kernel", [0, 1], ids=["mfma", "rows"])
def test_self_lists_match_oracle(m2l_kernel):
    # rinv(Q, U, is_self) with is_self = membership in self_generator (fmm.hpp:398-408,786-807)
    pos, q = W.clusters(3000, seed=48)
    rng = np.random.default_rng(5)
    lists = _self_lists(pos, rng)
    part = api.OctreePartitioner(pos, api.Box(pos)).refine(20, 24)
    fmm = api.FMM3D(part, 6, ncharge=3, m2l_kernel=m2l_kernel, self_lists=lists)
    ot = O.Tree(pos, 20, 24)
    for which, owhich in ((api.BOTH, O.BOTH), (api.NEAR, O.NEAR)):
        U = np.zeros((3000, 4))
        fmm.rinv_potgrad(q, U, which | api.SELF_LIST)
        assert rel_l2(U, ot.fmm(6, q, O.POTGRAD, which=owhich, self_lists=lists)) < PARITY
    Q3 = rng.uniform(-1, 1, (3000, 3))
    F = np.zeros((3000, 3))
    fmm.rinv(Q3, F, gradient=True, which=api.BOTH | api.SELF_LIST)
    assert rel_l2(F, ot.fmm(6, Q3, O.CURL, self_lists=lists)) < PARITY


def test_self_lists_default_predicate_subtracts_far_pairs_only():
    # rinv(Q, U) on a plan built with a self generator: the near field skips j == i
    # only, and the generator's far pairs (_slist) are subtracted
    pos, q = W.uniform_cube(2000, seed=49)
    rng = np.random.default_rng(6)
    row, val = _self_lists(pos, rng)
    part = api.OctreePartitioner(pos, api.Box(pos)).refine(20, 16)
    ot = O.Tree(pos, 20, 16)
    with_lists = api.FMM3D(part, 5, ncharge=1, self_lists=(row, val))
    plain = api.FMM3D(part, 5, ncharge=1)
    a = np.zeros((2000, 4))
    b = np.zeros((2000, 4))
    with_lists.rinv_potgrad(q, a)
    plain.rinv_potgrad(q, b)
    expect = np.zeros((2000, 4))
    n_far = 0
    for i in range(2000):
        for j in val[row[i]:row[i + 1]]:
            if j == i or ot.is_neighbor(i, j):
                continue
            n_far += 1
            d = pos[i] - pos[j]
            r = np.linalg.norm(d)
            expect[i, 0] -= q[j] / r
            expect[i, 1:] += q[j] * d / r ** 3
    assert n_far > 1000
    assert np.max(np.abs((a - b) - expect)) < 1e-12 * np.max(np.abs(b))
