"""Pins the oracle's ConnectivityMap restatement (test infrastructure).

  * cell-list rows == brute-force O(N^2) rows with the exact criterion, on
    random positions with random anisotropic, rotated H in 1/2/3-D (the setup of
    src/Neighbor/tests/testNestedGridNeighbor.py:37-187, but checking equality
    with the exact criterion rather than a superset);
  * ConnectivityMap::valid() invariants (src/Neighbor/ConnectivityMap.cc:315-516):
    no self, rows sorted, symmetric between internal nodes;
  * lattice closed form: at nPerh = 1.51 and extent 2 the cutoff is 3.02 dx, so
    an interior node of a perfect lattice has 122 (3-D), 28 (2-D), 6 (1-D)
    neighbours (count of integer vectors with |n| <= 3.02).
"""
import numpy as np
import pytest

from sphgpu import problems


def random_H(dim, n, rng, hmin=0.05, hmax=0.15):
    """SymTensor H = R diag(1/h) R^T with random rotation and anisotropy."""
    ns = dim * (dim + 1) // 2
    H = np.zeros((n, ns))
    for i in range(n):
        h = rng.uniform(hmin, hmax, dim)
        Q, _ = np.linalg.qr(rng.standard_normal((dim, dim)))
        M = Q @ np.diag(1.0 / h) @ Q.T
        k = 0
        for a in range(dim):
            for b in range(a, dim):
                H[i, k] = 0.5 * (M[a, b] + M[b, a])
                k += 1
    return H


def check_valid(off, nbr, n_int):
    rows = [nbr[off[i]:off[i + 1]] for i in range(n_int)]
    for i, r in enumerate(rows):
        assert i not in r
        assert np.all(np.diff(r) > 0)
    sets = [set(r.tolist()) for r in rows]
    for i, s in enumerate(sets):
        for j in s:
            if j < n_int:
                assert i in sets[j]


@pytest.mark.parametrize("dim,n", [(1, 300), (2, 500), (3, 700)])
def test_cell_list_equals_brute_force(oracle, dim, n):
    rng = np.random.default_rng(100 + dim)
    pos = rng.uniform(0, 1, (n, dim))
    H = random_H(dim, n, rng)
    n_int = n - n // 7          # the rest act as ghosts: no rows, but can be neighbours
    off_c, nbr_c = oracle.connectivity(dim, n_int, pos, H, 2.0, brute=False)
    off_b, nbr_b = oracle.connectivity(dim, n_int, pos, H, 2.0, brute=True)
    assert np.array_equal(off_c, off_b)
    assert np.array_equal(nbr_c, nbr_b)
    check_valid(off_c, nbr_c, n_int)
    # exact criterion on a few pairs by hand
    def Hmat(h):
        M = np.zeros((dim, dim))
        k = 0
        for a in range(dim):
            for b in range(a, dim):
                M[a, b] = M[b, a] = h[k]
                k += 1
        return M
    for i in range(0, n_int, max(1, n_int // 20)):
        row = set(nbr_c[off_c[i]:off_c[i + 1]].tolist())
        for j in range(n):
            if j == i:
                continue
            r = pos[i] - pos[j]
            e2i = np.sum((Hmat(H[i]) @ r) ** 2)
            e2j = np.sum((Hmat(H[j]) @ r) ** 2)
            margin = 1e-9
            if min(e2i, e2j) < 4.0 * (1 - margin):
                assert j in row
            elif min(e2i, e2j) > 4.0 * (1 + margin):
                assert j not in row


@pytest.mark.parametrize("dim,n,expect", [(1, 40, 6), (2, 24, 28), (3, 14, 122)])
def test_lattice_neighbor_count(oracle, dim, n, expect):
    pos, mass, H, dx = problems.lattice(dim, n, jitter=0.0, nPerh=1.51)
    N = pos.shape[0]
    off, nbr = oracle.connectivity(dim, N, pos, H, 2.0)
    counts = np.diff(off)
    # closed form: integer vectors with 0 < |m| <= 3.02
    r = np.arange(-4, 5)
    grids = np.meshgrid(*([r] * dim), indexing="ij")
    m2 = sum(g.astype(float) ** 2 for g in grids)
    closed = int(np.sum((m2 > 0) & (m2 <= 3.02 ** 2)))
    assert closed == expect
    interior = np.all((pos > 3.5 * dx) & (pos < 1 - 3.5 * dx), axis=1)
    assert interior.any()
    assert np.all(counts[interior] == expect)
    assert np.all(counts <= expect)
    check_valid(off, nbr, N)
