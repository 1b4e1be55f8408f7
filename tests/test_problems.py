"""Synthetic inputs: sizes of the benchmark configurations, structure of the
operators and determinism of the portable generator.  No GPU needed."""
import numpy as np
import scipy.sparse as sp


def test_splitmix_is_deterministic_and_uniform(dh):
    P = dh.problems
    u = P.splitmix_uniform(20261202, 100000)
    assert np.array_equal(u, P.splitmix_uniform(20261202, 100000))
    assert np.array_equal(u[10:20], P.splitmix_uniform(20261202, 10, offset=10))
    assert 0.0 <= u.min() and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.01
    z = P.splitmix_normal(7, 200001)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01
    # the first SplitMix64 output for state 0 is the published test value
    assert (np.uint64(0xE220A8397B1DCDAF) >> np.uint64(11)) * (1.0 / 9007199254740992.0) == \
        P.splitmix_uniform(0, 1)[0]


def test_config_sizes(dh):
    P = dh.problems
    assert P.poisson2d(256).nnz == 326656
    for m in (16, 32):
        assert P.heterogeneous3d(m, np.zeros(m ** 3)).nnz == 7 * m ** 3 - 6 * m ** 2
    # nnz of the 7-point operator on m^3 cells: 7 m^3 - 6 m^2
    for m in (128, 192, 96):
        assert 7 * m ** 3 - 6 * m ** 2 == {128: 14581760, 192: 49324032, 96: 6137856}[m]


def test_heterogeneous_operator_structure(dh):
    P = dh.problems
    m = 10
    logk = P.log_permeability(m, 1.5, 1)
    A = P.heterogeneous3d(m, logk)
    S = sp.csr_matrix((A.nzval, A.colval, A.rowptr), shape=(A.n, A.n))
    assert abs(S - S.T).max() == 0.0                       # symmetric
    assert np.all(np.diff(A.colval[A.rowptr[0]:A.rowptr[1]]) > 0)
    off = S - sp.diags(S.diagonal())
    assert off.max() <= 0.0                                # M-matrix sign pattern
    rowsum = np.asarray(S.sum(axis=1)).ravel()
    assert rowsum.min() >= -1e-9 * S.diagonal().max()      # weakly diagonally dominant
    # contrast of the cell permeabilities ~ 10^(+-3 sigma)
    k = np.exp(logk)
    assert k.max() / k.min() > 1e5
    # one face: T = 2 ka kb / (ka + kb) between cells 0 and 1
    ka, kb = k[0], k[1]
    assert np.isclose(-S[0, 1], 2 * ka * kb / (ka + kb), rtol=1e-15)
    # a corner cell has 3 Dirichlet faces: diag = sum T + 3 * 2 k
    T = [-S[0, 1], -S[0, m], -S[0, m * m]]
    assert np.isclose(S[0, 0], sum(T) + 6 * k[0], rtol=1e-14)


def test_face_list(dh):
    a, b = dh.problems.heterogeneous3d_faces(5)
    assert len(a) == 3 * 5 * 5 * 4 and np.all(a < b)
