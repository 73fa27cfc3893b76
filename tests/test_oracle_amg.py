"""Oracle AMG (oracle/amg.c) pinned against known answers and closed forms.

Reference: mtj/src/no/uib/cipr/matrix/sparse/AMG.java, SSOR.java; the HashSet
iteration order of the aggregation's phase 2 is Java's HashMap order (bucket
(h ^ h>>>16) & (cap-1), insertion order inside a bucket).
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from oracle import pyoracle as po
from ressim_amd import problems as P
from tests.helpers import oracle_system


def c1(nx, ny):
    """2-D MPFA-O pressure matrix (config-1 generator at nx x ny) with the
    centre cell locked: (rowptr, colind, data, b)."""
    d = oracle_system(po, P.config1(nx=nx, ny=ny))
    return d["rowptr"], d["colind"], d["data"], d["b"]


def csr(A):
    A = sp.csr_matrix(A)
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64)


def lap1d(n):
    return sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1]).tocsr()


# ---------------------------------------------------------------- HashSet order
@pytest.mark.parametrize("cols,expect", [
    ([1, 2, 3], [1, 2, 3]),            # small Integers iterate ascending
    ([1, 16], [16, 1]),                 # 16 lands in bucket 0 of a 16-table
    ([1, 17], [1, 17]),                 # same bucket: insertion order
    ([2, 65537], [65537, 2]),           # 65537 ^ (65537 >>> 16) = 65536 -> bucket 0
    # 12 entries stay in a 16-table (16 -> bucket 0); the 13th triggers the
    # resize to 32 slots (threshold 0.75 * 16), where 16 has its own bucket
    (list(range(11)) + [16], [0, 16] + list(range(1, 11))),
    (list(range(12)) + [16], list(range(12)) + [16]),
    ([30, 31, 35], [35, 30, 31]),
])
def test_hashset_order(cols, expect):
    assert list(po.hashset_order(np.array(cols, np.int32))) == expect


# ---------------------------------------------------------------- aggregation
def test_aggregates_1d_greedy():
    """Tridiagonal: N(i) = {i-1, i, i+1}; the greedy phase 1 takes {0,1}, then
    every third node's neighbourhood."""
    rp, ci, v = csr(lap1d(10))
    pt, nagg, nroots = po.amg_aggregate(rp, ci, v, 0.08, mode=0)
    assert list(pt) == [0, 0, 1, 1, 1, 2, 2, 2, 3, 3]
    assert nagg == 4 and nroots == 4


def tie_graph():
    """Two clusters joined through node 35; node 31 overlaps aggregate 0
    (via 30) and aggregate 1 (via 35) once each.  Other nodes are isolated."""
    n = 40
    A = sp.lil_matrix((n, n))
    for i in range(n):
        A[i, i] = 2.0
    for a, b in [(29, 30), (30, 31), (31, 35), (35, 32), (32, 33), (33, 34)]:
        A[a, b] = A[b, a] = -1.0
    return csr(A)


def test_aggregates_phase2_hashset_tie():
    rp, ci, v = tie_graph()
    pt, nagg, nroots = po.amg_aggregate(rp, ci, v, 0.08, mode=0)
    assert nroots == 2 and nagg == 2
    assert pt[29] == pt[30] == 0
    assert pt[32] == pt[33] == pt[35] == 1 and pt[34] == 1
    # N(31) = {30, 31, 35} iterates as [35, 30, 31]: aggregate 1 reaches the
    # largest overlap first (ascending order would pick aggregate 0)
    assert pt[31] == 1
    iso = [i for i in range(40) if i not in (29, 30, 31, 32, 33, 34, 35)]
    assert all(pt[i] == -1 for i in iso)


def _strength_sets(rp, ci, v, eps):
    A = sp.csr_matrix((v, ci, rp))
    d = A.diagonal()
    N = []
    for i in range(A.shape[0]):
        cols = ci[rp[i]:rp[i + 1]]
        vals = v[rp[i]:rp[i + 1]]
        N.append({int(c) for c, a in zip(cols, vals) if abs(a) >= eps * np.sqrt(d[i] * d[c])})
    return N


@pytest.mark.parametrize("mode", [0, 1])
def test_aggregates_are_partition(mode):
    rp, ci, data, _ = c1(24, 20)
    pt, nagg, nroots = po.amg_aggregate(rp, ci, data, 0.08, mode=mode)
    N = _strength_sets(rp, ci, data, 0.08)
    A = sp.csr_matrix((data, ci, rp))
    offd = (abs(A - sp.diags(A.diagonal())) > 0).sum(axis=1).A.ravel() > 0
    # every connected node is aggregated, isolated (pinned) rows are not
    assert np.all((pt >= 0) == offd)
    assert set(pt[pt >= 0]) == set(range(nagg))
    # each aggregate is connected through strong couplings
    for a in range(nagg):
        nodes = set(np.flatnonzero(pt == a).tolist())
        seen, stack = set(), [next(iter(nodes))]
        while stack:
            u = stack.pop()
            if u in seen:
                continue
            seen.add(u)
            stack += [w for w in N[u] | {j for j in nodes if u in N[j]} if w in nodes]
        assert seen == nodes
    # initial aggregates are whole, disjoint strong neighbourhoods
    roots = [i for i in range(len(pt)) if offd[i] and N[i] <= set(np.flatnonzero(pt == pt[i]))]
    assert nroots <= nagg and len(roots) >= nroots


# ---------------------------------------------------------------- prolongator / Galerkin
def _dense_P(A, pt, nagg, eps, omega):
    """(I - omega D^-1 A_F) P_tent with weak couplings lumped into A_F's
    diagonal (Vanek-Mandel-Brezina smoothed aggregation)."""
    A = A.toarray()
    n = A.shape[0]
    d = np.diag(A).copy()
    AF = A.copy()
    for i in range(n):
        if pt[i] < 0:
            continue
        for j in range(n):
            if j == i or A[i, j] == 0 or pt[j] < 0:
                continue
            if abs(A[i, j]) < eps * np.sqrt(d[i] * d[j]):
                AF[i, i] -= A[i, j]
                AF[i, j] = 0
    T = np.zeros((n, nagg))
    for i in range(n):
        if pt[i] >= 0:
            T[i, pt[i]] = 1
    for i in range(n):
        if pt[i] < 0:
            AF[i, :] = 0
    AF[:, pt < 0] = 0
    P = T - omega * (AF / d[:, None]) @ T
    P[pt < 0, :] = 0
    return P


@pytest.mark.parametrize("case", ["lap1d", "c1", "aniso"])
def test_prolongator_and_galerkin_closed_form(case):
    if case == "lap1d":
        rp, ci, v = csr(lap1d(60))
    elif case == "c1":
        rp, ci, v, _ = c1(20, 18)
    else:
        # anisotropic 2-D: weak y couplings get lumped
        n1 = 14
        I = sp.identity(n1)
        A = sp.kron(I, lap1d(n1)) + 0.01 * sp.kron(lap1d(n1), I)
        rp, ci, v = csr(A)
    H = po.AMG(rp, ci, v, min=8)
    lv = H.levels()
    assert len(lv) >= 2
    A = sp.csr_matrix((v, ci, rp))
    for k in range(len(lv) - 1):
        rpk, cik, vk, pt, prp, pci, pv = H.level(k)
        Ak = sp.csr_matrix((vk, cik, rpk), shape=(lv[k][0], lv[k][0]))
        nagg = lv[k + 1][0]
        P = sp.csr_matrix((pv, pci, prp), shape=(lv[k][0], nagg)).toarray()
        Pd = _dense_P(Ak, pt, nagg, 0.08 * 0.5 ** k, 2. / 3)
        np.testing.assert_allclose(P, Pd, rtol=1e-13, atol=1e-15)
        rpc, cic, vc, *_ = H.level(k + 1)
        Ac = sp.csr_matrix((vc, cic, rpc), shape=(nagg, nagg)).toarray()
        ref = P.T @ Ak.toarray() @ P
        scale = abs(ref).max()
        np.testing.assert_allclose(Ac, ref, rtol=0, atol=1e-13 * scale)
        # only nonzero entries are stored, rows sorted
        assert np.all(vc != 0)
        for i in range(nagg):
            assert np.all(np.diff(cic[rpc[i]:rpc[i + 1]]) > 0)


def test_single_level_is_exact_solve():
    """n <= min: the hierarchy is the dense LU of A, so apply = A^-1 b."""
    rp, ci, v, _ = c1(6, 5)
    A = sp.csr_matrix((v, ci, rp)).toarray()
    H = po.AMG(rp, ci, v, min=1000)
    assert len(H.levels()) == 1
    rng = np.random.default_rng(3)
    b = rng.standard_normal(A.shape[0])
    x = H.apply(b, rng.standard_normal(A.shape[0]))
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-12, atol=1e-12)
    xt = H.apply(b, np.zeros_like(b), transpose=True)
    np.testing.assert_allclose(xt, np.linalg.solve(A.T, b), rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------- SSOR
def _ssor_matrix_form(A, b, x, wf, wr, reverse, blocks):
    """Matrix form of one (block) SSOR step: in-block strict lower/upper parts
    L_b, U_b, off-block part from the starting vector."""
    n = A.shape[0]
    blk = np.zeros(n, int)
    for k, (s, e) in enumerate(blocks):
        blk[s:e] = k
    Ad = A.toarray()
    D = np.diag(np.diag(Ad))
    same = blk[:, None] == blk[None, :]
    Lb = np.tril(Ad * same, -1)
    Ub = np.triu(Ad * same, 1)
    Aoff = Ad * (~same)
    rhs = b - Aoff @ x
    xh = np.linalg.solve(D + wf * Lb, wf * rhs + ((1 - wf) * D - wf * Ub) @ x)
    if not reverse:
        return xh
    return np.linalg.solve(D + wr * Ub, wr * rhs + ((1 - wr) * D - wr * Lb) @ xh)


@pytest.mark.parametrize("block_rows", [0, 7, 16])
@pytest.mark.parametrize("reverse", [True, False])
@pytest.mark.parametrize("ordering", [0, 1])
def test_ssor_matrix_form(block_rows, reverse, ordering):
    """Natural order: SSOR.java's recurrences; colour order: SSOR of the
    symmetrically permuted matrix (visiting order = permutation)."""
    rp, ci, v, _ = c1(8, 6)
    A = sp.csr_matrix((v, ci, rp))
    n = A.shape[0]
    rng = np.random.default_rng(5)
    b, x = rng.standard_normal(n), rng.standard_normal(n)
    y, order = po.ssor_apply(rp, ci, v, b, x, omega_f=1.3, omega_r=0.8, reverse=reverse,
                             block_rows=block_rows, ordering=ordering, return_order=True)
    nb = 1 if block_rows == 0 else -(-n // block_rows)
    bounds = [(n * k // nb, n * (k + 1) // nb) for k in range(nb)]
    # visiting order is a permutation within each block
    for s0, e0 in bounds:
        assert sorted(order[s0:e0]) == list(range(s0, e0))
    if ordering == 0:
        assert list(order) == list(range(n))
    else:
        assert not np.array_equal(order, np.arange(n))
    perm = np.asarray(order)
    Ap = A[perm][:, perm]
    ref = _ssor_matrix_form(Ap, b[perm], x[perm], 1.3, 0.8, reverse, bounds)
    yr = np.empty(n)
    yr[perm] = ref
    np.testing.assert_allclose(y, yr, rtol=1e-12, atol=1e-12)


def test_multicolour_order_is_a_colouring():
    """Greedy colour classes: consecutive runs of mutually uncoupled rows
    whose count matches the 9-point stencil's 4-colouring in 2-D."""
    rp, ci, v, _ = c1(10, 9)
    A = sp.csr_matrix((v, ci, rp))
    n = A.shape[0]
    _, order = po.ssor_apply(rp, ci, v, np.zeros(n), np.zeros(n), ordering=1,
                             return_order=True)
    S = ((abs(A) + abs(A.T)) > 0).toarray()
    # split the visiting order into maximal runs of pairwise uncoupled rows
    runs, cur = [], []
    for r in order:
        if any(S[r, q] for q in cur):
            runs.append(cur)
            cur = []
        cur.append(r)
    runs.append(cur)
    assert len(runs) == 4
    for run in runs:
        assert not S[np.ix_(run, run)][~np.eye(len(run), dtype=bool)].any()


# ---------------------------------------------------------------- as a preconditioner
def test_amg_zero_guess_is_linear_and_converges():
    rp, ci, v, b = c1(32, 32)
    A = sp.csr_matrix((v, ci, rp))
    H = po.AMG(rp, ci, v, zero_guess=1)
    rng = np.random.default_rng(1)
    r1, r2 = rng.standard_normal(A.shape[0]), rng.standard_normal(A.shape[0])
    y = H.apply(2.0 * r1 - 3.0 * r2, rng.standard_normal(A.shape[0]))
    y12 = 2.0 * H.apply(r1, np.zeros_like(r1)) - 3.0 * H.apply(r2, np.zeros_like(r2))
    np.testing.assert_allclose(y, y12, rtol=1e-10, atol=1e-12)
    x, st, _ = po.solve("gmres", rp, ci, v, b, rtol=1e-10, prec="amg",
                        amg=dict(zero_guess=1), max_iter=400)
    assert st.status == 0
    xe = spl.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - xe) / np.linalg.norm(xe) < 1e-8
    # far fewer iterations than ILU(0)
    _, sti, _ = po.solve("gmres", rp, ci, v, b, rtol=1e-10, prec="ilu", max_iter=2000)
    assert st.iterations * 2 < sti.iterations


def test_amg_cycle_reduces_error():
    """A V-cycle from a nonzero start reduces the error of A x = b."""
    rp, ci, v, b = c1(32, 32)
    A = sp.csr_matrix((v, ci, rp))
    xe = spl.spsolve(A.tocsc(), b)
    H = po.AMG(rp, ci, v)
    x = np.zeros_like(b)
    errs = []
    for _ in range(6):
        # stationary iteration x <- x + B (b - A x) through the x-start form:
        # B applied to b with start x is x + B(b - A x) for the linear cycle
        x = H.apply(b, x)
        errs.append(np.linalg.norm(x - xe) / np.linalg.norm(xe))
    assert errs[-1] < 0.5 * errs[0] and all(e2 < e1 for e1, e2 in zip(errs, errs[1:]))
