"""The product's host setup (C++, draugr.jl_amd/csrc/host) against the
oracle's pure-Python restatements (oracle/setup_ref.py) bit for bit, and
against the setup invariants the reference tests assert.  No GPU needed."""
import numpy as np
import pytest

from oracle import setup_ref as ref

SEED = 20261202


def _anisotropic(dh, n, kx=1e4, ky=1e-2):
    """test/runtests.jl:1681-1694 style anisotropic 5-point operator."""
    A = dh.problems.poisson2d(n)
    rows = np.repeat(np.arange(A.n), np.diff(A.rowptr))
    d = np.abs(A.colval - rows)
    nz = np.where(d == 1, -kx, np.where(d == n, -ky, 0.0))
    diag = np.zeros(A.n)
    np.add.at(diag, rows, -nz)
    nz = np.where(d == 0, diag[rows] + 1e-3, nz)
    return dh.problems.CSR(A.rowptr, A.colval, nz)


def _random_spd(dh, n, seed):
    rng = np.random.default_rng(seed)
    import scipy.sparse as sp
    M = sp.random(n, n, density=4.0 / n, random_state=seed, format="csr")
    M = abs(M + M.T)
    M.setdiag(0)
    M.eliminate_zeros()
    A = -M
    A = A + sp.diags(np.asarray(M.sum(axis=1)).ravel() + 0.1 + rng.random(n))
    A = A.tocsr()
    A.sort_indices()
    return dh.problems.CSR(A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy())


def _mats(dh):
    P = dh.problems
    return {"poisson1d": P.poisson1d(30), "poisson2d": P.poisson2d(13), "aniso": _anisotropic(dh, 11),
            "random": _random_spd(dh, 300, SEED), "hetero3d": P.heterogeneous3d(6, P.log_permeability(6, 1.5, 3))}


@pytest.mark.parametrize("name", ["poisson1d", "poisson2d", "aniso", "random", "hetero3d"])
def test_aggregation_matches_restatement(dh, name):
    A = _mats(dh)[name]
    for theta in (0.25, 0.08):
        cfg = dh.AMGConfig(coarsening=dh.AggregationCoarsening(theta), smoother=dh.JacobiSmootherType(),
                           max_coarse_size=4)
        h = dh.amg_setup(A, cfg, upload=False)
        lv = h.level(0)
        agg, nc = ref.coarsen_aggregation(A.rowptr, A.colval, A.nzval, theta)
        assert lv["n_coarse"] == nc
        assert np.array_equal(lv["P_cv"], np.asarray(agg, np.int32))
        assert np.array_equal(lv["P_nz"], np.ones(A.n))


@pytest.mark.parametrize("name", ["poisson2d", "aniso", "random", "hetero3d"])
@pytest.mark.parametrize("config", ["aggregation", "hypre", "pmis_direct", "rs_standard"])
def test_coarse_sparsity_transpose_map_match_restatement(dh, name, config):
    """Coarse pattern, values (same summation order), triple map and transpose
    map of every level equal the restatement of src/galerkin.jl:11-183 and
    src/prolongation.jl:767-797 exactly."""
    A = _mats(dh)[name]
    cfg = {"aggregation": dh.AMGConfig(coarsening=dh.AggregationCoarsening(0.25), smoother=dh.JacobiSmootherType(),
                                       max_coarse_size=8),
           "hypre": dh.hypre_default_config(smoother=dh.L1JacobiSmootherType(), max_coarse_size=8),
           "pmis_direct": dh.AMGConfig(coarsening=dh.PMISCoarsening(0.25), smoother=dh.JacobiSmootherType(),
                                       max_coarse_size=8),
           "rs_standard": dh.AMGConfig(coarsening=dh.RSCoarsening(0.25, dh.StandardInterpolation(0.2)),
                                       smoother=dh.JacobiSmootherType(), max_coarse_size=8)}[config]
    h = dh.amg_setup(A, cfg, upload=False)
    for l in range(h.nlevels):
        lv = h.level(l)
        Arp, Acv, Anz = lv["A_rp"], lv["A_cv"], lv["A_nz"]
        P = (lv["P_rp"], lv["P_cv"], lv["P_nz"])
        (rp, cv, vals), rm = ref.coarse_sparsity((Arp, Acv, Anz), P, lv["n_coarse"])
        if l + 1 < h.nlevels:
            nxt = h.level(l + 1)
            grp, gcv, gnz = nxt["A_rp"], nxt["A_cv"], nxt["A_nz"]
        else:
            c = h.coarsest()
            grp, gcv, gnz = c["rp"], c["cv"], c["nz"]
        assert np.array_equal(grp, rp) and np.array_equal(gcv, cv)
        assert np.array_equal(gnz, vals)
        for k in ("R_off", "R_pi", "R_ai", "R_pj"):
            assert np.array_equal(lv[k], rm[k]), k
        off, rows, pnz = ref.transpose_map(P, lv["n_coarse"])
        assert np.array_equal(lv["Pt_off"], off) and np.array_equal(lv["Pt_rows"], rows)
        assert np.array_equal(lv["Pt_pnz"], pnz)


@pytest.mark.parametrize("name", ["poisson2d", "random", "hetero3d"])
def test_coloring_matches_restatement_and_is_valid(dh, name):
    A = _mats(dh)[name]
    cfg = dh.AMGConfig(coarsening=dh.AggregationCoarsening(0.25), smoother=dh.ColoredGaussSeidelType(),
                       max_coarse_size=8)
    h = dh.amg_setup(A, cfg, upload=False)
    for l in range(h.nlevels):
        lv = h.level(l)
        colors, off, order = ref.greedy_coloring(lv["A_rp"], lv["A_cv"])
        assert np.array_equal(lv["c_off"], off) and np.array_equal(lv["c_order"], order)
        col = np.empty(lv["n"], np.int32)
        for c in range(len(off) - 1):
            col[order[off[c]:off[c + 1]]] = c
        rows = np.repeat(np.arange(lv["n"]), np.diff(lv["A_rp"]))
        offd = rows != lv["A_cv"]
        assert not np.any(col[rows[offd]] == col[lv["A_cv"][offd]])


def test_aggregation_prolongation_known_answers(dh, orc):
    """P*1 = 1 and sum(P^T r) = sum(r) for aggregation (test/runtests.jl:138-179)."""
    A = dh.problems.poisson1d(10)
    h = dh.amg_setup(A, dh.AMGConfig(coarsening=dh.AggregationCoarsening(), smoother=dh.JacobiSmootherType(),
                                     max_coarse_size=2), upload=False)
    lv = h.level(0)
    assert np.array_equal(np.diff(lv["P_rp"]), np.ones(10, np.int32))
    xf = orc.prolong(lv["P_rp"], lv["P_cv"], lv["P_nz"], np.ones(lv["n_coarse"]), np.zeros(10))
    assert np.array_equal(xf, np.ones(10))
    r = np.random.default_rng(SEED).standard_normal(10)
    bc = orc.restrict(lv["Pt_off"], lv["Pt_rows"], lv["Pt_pnz"], lv["P_nz"], r)
    assert abs(bc.sum() - r.sum()) < 1e-13
    assert lv["Pt_off"][-1] == 10


def test_hmis_hierarchy_shape(dh):
    """HMIS with ext+i on 2-D Poisson: <= 10 levels at 50x50 and a first
    coarsening ratio < 0.6 at 20x20 (test/runtests.jl:967-987)."""
    cfg = dh.AMGConfig(coarsening=dh.HMISCoarsening(0.5, dh.ExtendedIInterpolation()))
    h = dh.amg_setup(dh.problems.poisson2d(50), cfg, upload=False)
    assert h.nlevels + 1 <= 10
    h2 = dh.amg_setup(dh.problems.poisson2d(20), cfg, upload=False)
    assert h2.level(0)["n_coarse"] / 400 < 0.6


@pytest.mark.parametrize("config", ["hypre", "pmis_direct", "hmis_ext"])
def test_cf_interpolation_invariants(dh, config):
    """C rows of P are unit rows; every row of P has an entry; ext+i rows of a
    zero-row-sum interior sum to 1 (interpolation of constants)."""
    P_ = dh.problems
    A = P_.heterogeneous3d(8, np.zeros(512))  # homogeneous: interior rows sum to 0
    cfg = {"hypre": dh.hypre_default_config(),
           "pmis_direct": dh.AMGConfig(coarsening=dh.PMISCoarsening(0.25)),
           "hmis_ext": dh.AMGConfig(coarsening=dh.HMISCoarsening(0.25, dh.ExtendedIInterpolation()))}[config]
    h = dh.amg_setup(A, cfg, upload=False)
    lv = h.level(0)
    lens = np.diff(lv["P_rp"])
    assert np.all(lens >= 1)
    rows = np.repeat(np.arange(A.n), lens)
    rowsum = np.zeros(A.n)
    np.add.at(rowsum, rows, lv["P_nz"])
    Arows = np.repeat(np.arange(A.n), np.diff(A.rowptr))
    arow = np.zeros(A.n)
    np.add.at(arow, Arows, A.nzval)
    interior = np.abs(arow) < 1e-12
    assert np.allclose(rowsum[interior], 1.0, atol=1e-12)
    # unit rows: a row with a single entry 1.0 is a C point; there are n_coarse of them at least
    unit = (lens == 1) & np.isclose(lv["P_nz"][lv["P_rp"][:-1]], 1.0)
    assert unit.sum() >= lv["n_coarse"]


def test_setup_is_deterministic_for_a_seed(dh):
    A = dh.problems.heterogeneous3d(10, dh.problems.log_permeability(10, 1.5, 9))
    cfg = dh.hypre_default_config()
    a = dh.amg_setup(A, cfg, upload=False).level(0)
    b = dh.amg_setup(A, cfg, upload=False).level(0)
    for k in a:
        assert np.array_equal(a[k], b[k])
    c = dh.amg_setup(A, dh.hypre_default_config(seed=12345), upload=False).level(0)
    assert c["n_coarse"] > 0


def test_setup_generator_is_the_problems_splitmix(dh):
    """The oracle's stand-in for the reference's rand(rng) is the same
    SplitMix64 stream as the product's own counter-based generator."""
    for seed in (0x5EED, SEED, 1):
        assert np.array_equal(ref.splitmix_uniforms(seed, 257), dh.problems.splitmix_uniform(seed, 257))


@pytest.mark.parametrize("coarsening", ["pmis", "hmis", "aggressive_hmis", "aggressive_pmis"])
@pytest.mark.parametrize("name", ["poisson1d", "poisson2d", "aniso", "hetero3d", "random"])
def test_pmis_matches_reference_sequential_sweeps(dh, name, coarsening):
    """The host's PMIS, HMIS and aggressive CF coarsening (snapshot sweeps, run
    in parallel; HMIS's RS first pass over bucket lists; the second stage over
    the C points' distance-2 graph) decide every node as the reference's
    sequential code does (src/coarsening.jl:202-283, 291-313, 330-624,
    990-1116), on the setup's seeded draws: with direct interpolation each C point's row of P is its
    coarse index with weight 1, so every C point of the restatement must have
    exactly that row, and the coarse counts agree."""
    A = _mats(dh)[name]
    rp, cv, nz = (np.asarray(a) for a in (A.rowptr, A.colval, A.nzval))
    theta = 0.25
    alg = {"pmis": dh.PMISCoarsening(theta), "hmis": dh.HMISCoarsening(theta),
           "aggressive_hmis": dh.AggressiveCoarsening(theta, "hmis", dh.DirectInterpolation()),
           "aggressive_pmis": dh.AggressiveCoarsening(theta, "pmis", dh.DirectInterpolation())}[coarsening]
    cfg = dh.AMGConfig(coarsening=alg, smoother=dh.JacobiSmootherType(), max_coarse_size=4)
    h = dh.amg_setup(A, cfg, upload=False)
    if coarsening.startswith("aggressive"):
        cf = ref.coarsen_aggressive_cf(rp, cv, nz, theta, coarsening.endswith("hmis"), cfg.seed)
    else:
        cf = (ref.coarsen_pmis if coarsening == "pmis" else ref.coarsen_hmis)(rp, cv, nz, theta, cfg.seed)
    n = len(rp) - 1
    L = h.level(0, with_rmap=False)
    assert L["n_coarse"] < 0.8 * n  # no theta-halving retry (which would draw again)
    assert L["n_coarse"] == int((cf == 1).sum())
    coarse = np.cumsum(cf == 1) - 1
    Prp, Pcv, Pnz = L["P_rp"], L["P_cv"], L["P_nz"]
    for i in np.flatnonzero(cf == 1):
        assert Prp[i + 1] - Prp[i] == 1 and Pcv[Prp[i]] == coarse[i] and Pnz[Prp[i]] == 1.0, (name, i)


@pytest.mark.parametrize("interp", ["plain", "trunc", "trunc_max_rescale", "norm2"])
@pytest.mark.parametrize("name", ["poisson2d", "aniso", "hetero3d", "random"])
def test_ext_i_matches_reference(dh, name, interp):
    """The host's extended+i interpolation with its truncations
    (src/prolongation.jl:467-690, then src/setup.jl:263-266) gives the
    reference's P bit for bit on the HMIS splitting pinned above: same rows,
    columns and values (the GPU build is pinned to the host's by
    tests/test_gpu_setup.py)."""
    A = _mats(dh)[name]
    rp, cv, nz = (np.asarray(a) for a in (A.rowptr, A.colval, A.nzval))
    theta = 0.25
    spec = {"plain": dict(trunc_factor=0.0), "trunc": dict(trunc_factor=0.3),
            "trunc_max_rescale": dict(trunc_factor=0.1, max_elements=3, rescale=True),
            "norm2": dict(trunc_factor=0.3, norm_p=2)}[interp]
    cfg = dh.AMGConfig(coarsening=dh.HMISCoarsening(theta, dh.ExtendedIInterpolation(**spec)),
                       smoother=dh.JacobiSmootherType(), max_coarse_size=4)
    h = dh.amg_setup(A, cfg, upload=False)
    L = h.level(0, with_rmap=False)
    assert L["n_coarse"] < 0.8 * (len(rp) - 1)
    cf = ref.coarsen_hmis(rp, cv, nz, theta, cfg.seed)
    P = ref.interp_ext_i(rp, cv, nz, cf, theta, spec.get("trunc_factor", 0.0), spec.get("max_elements", 0),
                         spec.get("norm_p", 1), spec.get("rescale", False))
    if spec["trunc_factor"] > 0:
        P = ref.filter_prolongation(*P, spec["trunc_factor"])
    assert np.array_equal(L["P_rp"], P[0]) and np.array_equal(L["P_cv"], P[1])
    assert L["P_nz"].tobytes() == P[2].tobytes()


@pytest.mark.parametrize("interp", ["direct_pmis", "direct_hmis", "standard_pmis", "standard_hmis", "direct_rs",
                                    "standard_rs"])
@pytest.mark.parametrize("name", ["poisson2d", "aniso", "hetero3d", "random"])
def test_direct_standard_interpolation_match_reference(dh, name, interp):
    """The host's direct and standard interpolations (src/prolongation.jl:219-335,
    349-456) give the reference's P bit for bit on the pinned PMIS / HMIS
    splittings and on classical RS (src/coarsening.jl:671-818), which this
    pins too."""
    A = _mats(dh)[name]
    rp, cv, nz = (np.asarray(a) for a in (A.rowptr, A.colval, A.nzval))
    theta = 0.25
    kind, base = interp.split("_")
    ip = dh.DirectInterpolation() if kind == "direct" else dh.StandardInterpolation()
    alg = {"pmis": dh.PMISCoarsening, "hmis": dh.HMISCoarsening, "rs": dh.RSCoarsening}[base](theta, ip)
    cfg = dh.AMGConfig(coarsening=alg, smoother=dh.JacobiSmootherType(), max_coarse_size=4)
    h = dh.amg_setup(A, cfg, upload=False)
    L = h.level(0, with_rmap=False)
    assert L["n_coarse"] < 0.8 * (len(rp) - 1)
    if base == "rs":
        cf = ref.coarsen_rs(rp, cv, nz, theta)
    else:
        cf = (ref.coarsen_pmis if base == "pmis" else ref.coarsen_hmis)(rp, cv, nz, theta, cfg.seed)
    P = (ref.interp_direct if kind == "direct" else ref.interp_standard)(rp, cv, nz, cf, theta)
    assert np.array_equal(L["P_rp"], P[0]) and np.array_equal(L["P_cv"], P[1])
    assert L["P_nz"].tobytes() == P[2].tobytes()


@pytest.mark.parametrize("filtering", [False, True])
@pytest.mark.parametrize("name", ["poisson1d", "poisson2d", "aniso", "random", "hetero3d"])
def test_smoothed_aggregation_matches_reference(dh, name, filtering):
    """The host's smoothed-aggregation P = (I - w D^-1 A) P_tent
    (src/prolongation.jl:38-120), and with filtering the entries below
    0.1 max |p| of each row dropped (:134-182, src/setup.jl:286-290), bit for
    bit on the pinned aggregation."""
    A = _mats(dh)[name]
    rp, cv, nz = (np.asarray(a) for a in (A.rowptr, A.colval, A.nzval))
    theta = 0.08
    alg = dh.SmoothedAggregationCoarsening(theta, filtering=filtering, filter_tol=0.1)
    cfg = dh.AMGConfig(coarsening=alg, smoother=dh.JacobiSmootherType(), max_coarse_size=4)
    h = dh.amg_setup(A, cfg, upload=False)
    L = h.level(0, with_rmap=False)
    agg, nc = ref.coarsen_aggregation(rp, cv, nz, theta)
    assert L["n_coarse"] == nc
    (prp, pcv, pnz), _ = ref.prolongation_from_aggregates(agg, nc)
    P = ref.smooth_prolongation(rp, cv, nz, prp, pcv, pnz, alg.ω)
    if filtering:
        P = ref.filter_prolongation(*P, 0.1)
    assert np.array_equal(L["P_rp"], P[0]) and np.array_equal(L["P_cv"], P[1])
    assert L["P_nz"].tobytes() == P[2].tobytes()


@pytest.mark.parametrize("name", ["poisson1d", "poisson2d", "aniso", "random", "hetero3d"])
def test_aggressive_aggregation_matches_reference(dh, name):
    """The host's aggressive coarsening on an aggregation base
    (src/coarsening.jl:869-957: PMIS, F points joined to their strongest C
    neighbour's aggregate, aggregates merged through F points) gives the
    reference's aggregates on the pinned PMIS splitting."""
    A = _mats(dh)[name]
    rp, cv, nz = (np.asarray(a) for a in (A.rowptr, A.colval, A.nzval))
    theta = 0.25
    cfg = dh.AMGConfig(coarsening=dh.AggressiveCoarsening(theta, "aggregation"), smoother=dh.JacobiSmootherType(),
                       max_coarse_size=4)
    h = dh.amg_setup(A, cfg, upload=False)
    L = h.level(0, with_rmap=False)
    agg, nc = ref.coarsen_aggressive_agg(rp, cv, nz, theta, cfg.seed)
    assert L["n_coarse"] == nc
    assert np.array_equal(L["P_cv"], agg) and np.array_equal(L["P_nz"], np.ones(len(rp) - 1))


@pytest.mark.parametrize("coarsening", ["hmis", "aggressive_hmis"])
@pytest.mark.parametrize("name", ["poisson2d", "random", "hetero3d"])  # (aniso's rows sum to ~1e-3: none weakened)
def test_max_row_sum_weakening_matches_reference(dh, name, coarsening):
    """With max_row_sum < 1 the C/F splitting sees the weakened matrix
    (src/strength.jl:183-213; src/coarsening.jl:999, 1120-1185) and the
    interpolation the original one: the host's P equals the restatement's."""
    A = _mats(dh)[name]
    rp, cv, nz = (np.asarray(a) for a in (A.rowptr, A.colval, A.nzval))
    theta, mrs = 0.25, 0.2  # boundary rows (|row sum| > 0.2 |a_ii|) are weakened
    ip = dh.DirectInterpolation()
    alg = dh.HMISCoarsening(theta, ip) if coarsening == "hmis" else dh.AggressiveCoarsening(theta, "hmis", ip)
    cfg = dh.AMGConfig(coarsening=alg, smoother=dh.JacobiSmootherType(), max_coarse_size=4, max_row_sum=mrs)
    h = dh.amg_setup(A, cfg, upload=False)
    L = h.level(0, with_rmap=False)
    assert L["n_coarse"] < 0.8 * (len(rp) - 1)
    weak = ref.apply_max_row_sum(rp, cv, nz, mrs)
    assert not np.array_equal(weak, nz)  # the weakening acts on these matrices
    if coarsening == "hmis":
        cf = ref.coarsen_hmis(rp, cv, weak, theta, cfg.seed)
    else:
        cf = ref.coarsen_aggressive_cf(rp, cv, weak, theta, True, cfg.seed)
    P = ref.interp_direct(rp, cv, nz, cf, theta)
    assert np.array_equal(L["P_rp"], P[0]) and np.array_equal(L["P_cv"], P[1])
    assert L["P_nz"].tobytes() == P[2].tobytes()
