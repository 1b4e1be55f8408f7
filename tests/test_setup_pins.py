"""Pins of the setup restatements (oracle/setup_ref.py) by what the
mathematics or the reference's own tests fix, independent of the product's
C++ host setup (which the restatements pin bit for bit elsewhere):

* strength of connection: the known answer of test/runtests.jl:96-102
  (1-D Poisson of 10 rows, theta 0.25: 18 strong entries);
* smoothed aggregation: P = (I - w D^-1 A) P_tent formed by scipy, entries of
  magnitude <= eps dropped (src/prolongation.jl:38-120);
* PMIS splittings: every F point with a strong connection strongly depends
  on a C point (src/coarsening.jl:255-264);
* direct interpolation: rows of zero row sum interpolate constants exactly
  (weights -a_ij / (a_ii + sum of the non-interpolatory off-diagonals)).

Each check is shown to reject a planted mistake (a mutated result or a
mutated formula).  No GPU needed."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import setup_ref as ref

SEED = 20261202
EPS = np.finfo(np.float64).eps


def _arr(A):
    return np.asarray(A.rowptr), np.asarray(A.colval), np.asarray(A.nzval)


def _mats(dh):
    P = dh.problems
    return {"poisson1d": P.poisson1d(30), "poisson2d": P.poisson2d(13),
            "hetero3d": P.heterogeneous3d(6, P.log_permeability(6, 1.5, 3)),
            "hetero3d_b": P.heterogeneous3d(7, P.log_permeability(7, 1.0, 11))}


# ── strength ────────────────────────────────────────────────────────────────
def test_strength_known_answer(dh):
    A = dh.problems.poisson1d(10)
    s = ref.strength_absolute(*_arr(A), 0.25)
    assert int(s.sum()) == 18
    # planted mistakes: the diagonal counted as a connection, or theta applied
    # to the diagonal's magnitude with a strict test (theta = 0.5: |-1| > 1)
    rp, cv, nz = _arr(A)
    diag_in = np.array([abs(nz[k]) >= 0.25 for i in range(A.n) for k in range(rp[i], rp[i + 1])])
    assert int(diag_in.sum()) != 18
    s_half = ref.strength_absolute(rp, cv, nz, 0.5)
    strict = np.array([(cv[k] != i) and abs(nz[k]) > 0.5 * 2.0 for i in range(A.n) for k in range(rp[i], rp[i + 1])])
    assert int(s_half.sum()) == 18 and int(strict.sum()) != 18


# ── smoothed aggregation ────────────────────────────────────────────────────
def _sa_scipy(A, P_tent, omega):
    As = sp.csr_matrix((A.nzval, A.colval, A.rowptr), shape=(A.n, A.n))
    d = As.diagonal()
    inv = np.where(np.abs(d) > EPS * np.maximum(1.0, np.abs(d)), 1.0 / np.where(d == 0, 1, d), 0.0)
    S = sp.identity(A.n, format="csr") - omega * sp.diags(inv) @ As
    P = (S @ P_tent).tocsr()
    P.sort_indices()
    return P


@pytest.mark.parametrize("name", ["poisson1d", "poisson2d", "hetero3d", "hetero3d_b"])
def test_smoothed_prolongation_equals_scipy_product(dh, name):
    A = _mats(dh)[name]
    rp, cv, nz = _arr(A)
    agg, nc = ref.coarsen_aggregation(rp, cv, nz, 0.08)
    (prp, pcv, pnz), _ = ref.prolongation_from_aggregates(agg, nc)
    omega = 2.0 / 3.0
    Pr = ref.smooth_prolongation(rp, cv, nz, prp, pcv, pnz, omega)
    Pt = sp.csr_matrix((pnz, pcv, prp), shape=(A.n, nc))
    Ps = _sa_scipy(A, Pt, omega)
    dense = Ps.toarray()
    keep = np.abs(dense) > EPS
    Pd = sp.csr_matrix((Pr[2], Pr[1], Pr[0]), shape=(A.n, nc)).toarray()
    assert np.array_equal(Pd != 0, keep)  # same pattern after the drop
    assert np.max(np.abs(Pd - np.where(keep, dense, 0.0))) < 1e-15
    # planted mistakes: the other sign of omega, or no drop of tiny entries
    Pbad = _sa_scipy(A, Pt, -omega).toarray()
    assert np.max(np.abs(Pd - Pbad)) > 1e-3
    tiny = (np.abs(dense) > 0) & (np.abs(dense) <= EPS)
    if tiny.any():
        assert not np.array_equal(Pd != 0, dense != 0)


# ── C/F splittings ──────────────────────────────────────────────────────────
def _f_without_strong_c(rp, cv, s, cf):
    """F points that strongly depend on some point but on no C point (an F
    point with no strong connection at all needs no interpolation: HMIS
    leaves such points F, src/coarsening.jl:372-381)."""
    bad = []
    for i in range(len(rp) - 1):
        if cf[i] != -1:
            continue
        strong = [cv[k] for k in range(rp[i], rp[i + 1]) if s[k] and cv[k] != i]
        if strong and not any(cf[j] == 1 for j in strong):
            bad.append(i)
    return bad


@pytest.mark.parametrize("name", ["poisson1d", "poisson2d", "hetero3d", "hetero3d_b"])
def test_every_pmis_f_point_has_a_strong_c_neighbour(dh, name):
    """PMIS only: HMIS's last stage makes points of measure < 1 F points
    whatever their neighbours (src/coarsening.jl:585-588), so the property is
    not one of HMIS."""
    A = _mats(dh)[name]
    rp, cv, nz = _arr(A)
    theta = 0.25
    cf = ref.coarsen_pmis(rp, cv, nz, theta, SEED)
    s = ref.strength_absolute(rp, cv, nz, theta)
    assert set(np.unique(cf)) <= {-1, 1} and (cf == 1).any() and (cf == -1).any()
    assert _f_without_strong_c(rp, cv, s, cf) == []
    # planted mistake: the strong C neighbours of one F point made F points
    i = int(np.flatnonzero(cf == -1)[0])
    bad = cf.copy()
    for k in range(rp[i], rp[i + 1]):
        if s[k] and cv[k] != i and cf[cv[k]] == 1:
            bad[cv[k]] = -1
    assert i in _f_without_strong_c(rp, cv, s, bad)


# ── direct interpolation ────────────────────────────────────────────────────
@pytest.mark.parametrize("name", ["poisson2d", "hetero3d", "hetero3d_b"])
def test_direct_interpolation_preserves_constants(dh, name):
    P = dh.problems
    if name == "poisson2d":
        A = P.poisson2d(13)
    else:  # homogeneous 3-D operator: every interior row sums to zero
        A = P.heterogeneous3d(6 if name == "hetero3d" else 7, np.zeros(216 if name == "hetero3d" else 343))
    rp, cv, nz = _arr(A)
    theta = 0.25
    cf = ref.coarsen_pmis(rp, cv, nz, theta, SEED)
    prp, pcv, pnz = ref.interp_direct(rp, cv, nz, cf, theta)
    rows = np.repeat(np.arange(A.n), np.diff(rp))
    arow = np.zeros(A.n)
    np.add.at(arow, rows, nz)
    zero_sum = np.abs(arow) < 1e-12
    assert zero_sum.sum() > A.n // 4
    psum = np.zeros(A.n)
    np.add.at(psum, np.repeat(np.arange(A.n), np.diff(prp)), pnz)
    assert np.max(np.abs(psum[zero_sum] - 1.0)) < 1e-14
    # planted mistake: a denominator without the non-interpolatory
    # off-diagonals (plain a_ii) no longer interpolates constants
    s = ref.strength_absolute(rp, cv, nz, theta)
    worst = 0.0
    for i in np.flatnonzero(zero_sum & (cf == -1)):
        a_ii = sum(nz[k] for k in range(rp[i], rp[i + 1]) if cv[k] == i)
        w = [-nz[k] / a_ii for k in range(rp[i], rp[i + 1])
             if cv[k] != i and s[k] and cf[cv[k]] == 1 and np.sign(nz[k]) != np.sign(a_ii)]
        if w:
            worst = max(worst, abs(sum(w) - 1.0))
    assert worst > 1e-3
