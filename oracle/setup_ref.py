"""ORACLE — test infrastructure only.

Pure-Python restatements of the setup pieces whose output the hot path
consumes as its contract, for small matrices (n up to a few thousand):

* ``coarsen_aggregation``   src/coarsening.jl:14-168 (deterministic)
* ``prolongation_from_aggregates`` src/prolongation.jl:8-22
* ``coarse_sparsity``       src/galerkin.jl:11-183 (pattern, values, triple map)
* ``transpose_map``         src/prolongation.jl:767-797
* ``greedy_coloring``       src/smoothers.jl:136-195
* ``coarsen_pmis``          src/coarsening.jl:202-283, with the setup's seeded
  SplitMix64 draws (``splitmix_uniforms``) in place of the reference's
  unseeded generator
* ``coarsen_hmis``          src/coarsening.jl:330-394 (``rs_first_pass``
  :406-567 with its bucket lists, ``pmis_on_undecided`` :575-624), same draws
* ``coarsen_aggressive_cf`` src/coarsening.jl:990-1116 (PMIS over the
  distance-2 graph of the first pass's C points) and
  ``ensure_fine_have_coarse_neighbor`` :291-313, same draws
* ``interp_ext_i``          src/prolongation.jl:467-690 (extended+i with its
  truncation, max_elements, norm_p and rescale) and ``filter_prolongation``
  :134-182, the second truncation of src/setup.jl:263-266
* ``interp_direct``, ``interp_standard`` src/prolongation.jl:219-335, 349-456
* ``coarsen_rs``            src/coarsening.jl:671-818 (classical first pass)
* ``smooth_prolongation``   src/prolongation.jl:38-120 (smoothed aggregation)
* ``coarsen_aggressive_agg`` src/coarsening.jl:869-957 (aggressive coarsening
  on an aggregation base: PMIS, aggregates merged through F points)
* ``apply_max_row_sum``     src/strength.jl:183-213 (the matrix the C/F
  splitting sees when max_row_sum < 1)

They pin the product's C++ setup bit for bit on the same input.  Indices are
0-based.  Matrices are (rp, cv, nz) numpy triples.
"""
from __future__ import annotations

import numpy as np


def strength_absolute(rp, cv, nz, theta):
    """src/strength.jl:24-61."""
    n = len(rp) - 1
    s = np.zeros(len(cv), dtype=bool)
    for i in range(n):
        mx = 0.0
        for k in range(rp[i], rp[i + 1]):
            if cv[k] != i:
                mx = max(mx, abs(nz[k]))
        thr = theta * mx
        for k in range(rp[i], rp[i + 1]):
            s[k] = (cv[k] != i) and abs(nz[k]) >= thr
    return s


def _find(parent, i):
    while parent[i] != i:
        parent[i] = parent[parent[i]]
        i = parent[i]
    return i


def coarsen_aggregation(rp, cv, nz, theta):
    """Distance-2 MIS seeds, neighbour assignment, singleton merge (0-based)."""
    n = len(rp) - 1
    if n <= 1:
        return [0] * n, n
    s = strength_absolute(rp, cv, nz, theta)
    agg = [-1] * n
    nc = 0
    scount = [0] * n
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            if cv[k] != i and s[k]:
                scount[i] += 1
    order = sorted(range(n), key=lambda i: -scount[i])  # stable
    state = [0] * n
    for i in order:
        if state[i] != 0:
            continue
        if any(cv[k] != i and s[k] and state[cv[k]] in (1, -1) for k in range(rp[i], rp[i + 1])):
            continue
        state[i] = 1
        agg[i] = nc
        nc += 1
        for k in range(rp[i], rp[i + 1]):
            j = cv[k]
            if j != i and s[k] and state[j] == 0:
                state[j] = -1
                agg[j] = agg[i]
                for k2 in range(rp[j], rp[j + 1]):
                    m = cv[k2]
                    if m != j and s[k2] and state[m] == 0:
                        state[m] = -2
    for i in range(n):
        if agg[i] >= 0:
            continue
        best, bv = -1, 0.0
        for k in range(rp[i], rp[i + 1]):
            j = cv[k]
            if j != i and s[k] and agg[j] >= 0 and abs(nz[k]) > bv:
                bv, best = abs(nz[k]), agg[j]
        if best < 0:
            for k in range(rp[i], rp[i + 1]):
                j = cv[k]
                if j != i and agg[j] >= 0 and abs(nz[k]) > bv:
                    bv, best = abs(nz[k]), agg[j]
        if best >= 0:
            agg[i] = best
        else:
            agg[i] = nc
            nc += 1
    size = [0] * nc
    for i in range(n):
        size[agg[i]] += 1
    parent = list(range(nc))
    for i in range(n):
        mine = _find(parent, agg[i])
        if size[mine] > 1:
            continue
        target, bv = -1, 0.0
        for k in range(rp[i], rp[i + 1]):
            j = cv[k]
            if j == i:
                continue
            rj = _find(parent, agg[j])
            if rj != mine and abs(nz[k]) > bv:
                bv, target = abs(nz[k]), rj
        if target >= 0:
            sm, stt = size[mine], size[target]
            ra, rb = _find(parent, target), _find(parent, mine)
            if ra != rb:
                parent[rb] = ra
            size[_find(parent, target)] = sm + stt
    new_id = [-1] * nc
    cnt = 0
    for a in range(nc):
        r = _find(parent, a)
        if new_id[r] < 0:
            new_id[r] = cnt
            cnt += 1
        new_id[a] = new_id[r]
    agg = [new_id[_find(parent, agg[i])] for i in range(n)]
    return agg, cnt


def prolongation_from_aggregates(agg, nc):
    n = len(agg)
    return (np.arange(n + 1, dtype=np.int32), np.asarray(agg, dtype=np.int32), np.ones(n)), nc


def coarse_sparsity(A, P, nc):
    """Pattern / values / triple map exactly as generated by the reference:
    triples in (i, P entry of i, A entry, P entry of j) order, grouped by
    coarse row, columns sorted and deduplicated, then grouped by nonzero with
    a stable sort (src/galerkin.jl:37-182)."""
    a_rp, a_cv, a_nz = A
    p_rp, p_cv, p_nz = P
    n = len(a_rp) - 1
    raw = []  # (I, J, pi, ai, pj)
    for i in range(n):
        for pi in range(p_rp[i], p_rp[i + 1]):
            I = p_cv[pi]
            for a in range(a_rp[i], a_rp[i + 1]):
                j = a_cv[a]
                for pj in range(p_rp[j], p_rp[j + 1]):
                    raw.append((I, p_cv[pj], pi, a, pj))
    rows = [[] for _ in range(nc)]
    for t, tr in enumerate(raw):
        rows[tr[0]].append(t)
    rp = [0]
    cvals = []
    for I in range(nc):
        cols = sorted(set(raw[t][1] for t in rows[I]))
        cvals.extend(cols)
        rp.append(len(cvals))
    rp = np.asarray(rp, dtype=np.int32)
    cvals = np.asarray(cvals, dtype=np.int32)
    vals = np.zeros(len(cvals))
    ci = np.zeros(len(raw), dtype=np.int64)
    for I in range(nc):
        lo, hi = rp[I], rp[I + 1]
        pos = {int(c): lo + k for k, c in enumerate(cvals[lo:hi])}
        for t in rows[I]:
            _, J, pi, a, pj = raw[t]
            k = pos[J]
            ci[t] = k
            vals[k] += p_nz[pi] * a_nz[a] * p_nz[pj]
    order = np.argsort(ci, kind="stable")
    off = np.zeros(len(cvals) + 1, dtype=np.int64)
    np.add.at(off, ci + 1, 1)
    off = np.cumsum(off)
    raw_a = np.asarray(raw, dtype=np.int64).reshape(-1, 5)
    return (rp, cvals, vals), {
        "R_off": off,
        "R_pi": raw_a[order, 2].astype(np.int32),
        "R_ai": raw_a[order, 3].astype(np.int32),
        "R_pj": raw_a[order, 4].astype(np.int32),
    }


def transpose_map(P, nc):
    p_rp, p_cv, _ = P
    n = len(p_rp) - 1
    off = np.zeros(nc + 1, dtype=np.int32)
    for k in range(len(p_cv)):
        off[p_cv[k] + 1] += 1
    off = np.cumsum(off).astype(np.int32)
    rows = np.zeros(len(p_cv), dtype=np.int32)
    pnz = np.zeros(len(p_cv), dtype=np.int32)
    pos = off[:-1].copy()
    for i in range(n):
        for k in range(p_rp[i], p_rp[i + 1]):
            J = p_cv[k]
            rows[pos[J]] = i
            pnz[pos[J]] = k
            pos[J] += 1
    return off, rows, pnz


def greedy_coloring(rp, cv):
    n = len(rp) - 1
    colors = [0] * n
    ncol = 0
    for i in range(n):
        used = {colors[cv[k]] for k in range(rp[i], rp[i + 1]) if cv[k] != i and colors[cv[k]] > 0}
        c = 1
        while c in used:
            c += 1
        colors[i] = c
        ncol = max(ncol, c)
    off = np.zeros(ncol + 1, dtype=np.int32)
    for c in colors:
        off[c] += 1
    off = np.cumsum(off).astype(np.int32)
    order = np.zeros(n, dtype=np.int32)
    pos = off[:-1].copy()
    for i in range(n):
        order[pos[colors[i] - 1]] = i
        pos[colors[i] - 1] += 1
    return np.asarray(colors, dtype=np.int32), off, order


_M64 = (1 << 64) - 1


def splitmix_uniforms(seed, n):
    """The setup's generator: SplitMix64 from ``seed``, each draw
    (z >> 11) * 2^-53 of the mixed counter (the k-th draw is a closed form
    of k).  Stands in for ``rand(rng)`` of src/coarsening.jl:218."""
    s = seed & _M64
    out = np.empty(n)
    for k in range(n):
        s = (s + 0x9E3779B97F4A7C15) & _M64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        z ^= z >> 31
        out[k] = (z >> 11) * (1.0 / 9007199254740992.0)
    return out


def strong_transpose_count(rp, cv, s):
    """src/coarsening.jl:179-192: rows that strongly depend on each node."""
    n = len(rp) - 1
    cnt = np.zeros(n, dtype=np.int64)
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            if s[k] and cv[k] != i:
                cnt[cv[k]] += 1
    return cnt


def coarsen_pmis(rp, cv, nz, theta, seed):
    """src/coarsening.jl:202-283, sweeping in place in row order as the
    reference does (1 = C, -1 = F); measure = strong column count + draw."""
    n = len(rp) - 1
    if n <= 1:
        return np.ones(n, dtype=np.int64)
    s = strength_absolute(rp, cv, nz, theta)
    st = strong_transpose_count(rp, cv, s)
    measure = st.astype(np.float64) + splitmix_uniforms(seed, n)
    cf = np.zeros(n, dtype=np.int64)
    for i in range(n):
        has = any(cv[k] != i and s[k] for k in range(rp[i], rp[i + 1]))
        if not has and st[i] == 0:
            cf[i] = 1
    for _ in range(n + 1):
        all_decided = True
        for i in range(n):
            if cf[i] != 0:
                continue
            all_decided = False
            is_max = True
            for k in range(rp[i], rp[i + 1]):
                j = cv[k]
                if s[k] and j != i and cf[j] != -1 and measure[j] > measure[i]:
                    is_max = False
                    break
            if is_max:
                cf[i] = 1
        for i in range(n):
            if cf[i] != 0:
                continue
            for k in range(rp[i], rp[i + 1]):
                if s[k] and cf[cv[k]] == 1:
                    cf[i] = -1
                    break
        if all_decided:
            break
    cf[cf == 0] = 1
    return cf


def _strong_transpose_adj(rp, cv, s):
    """src/coarsening.jl:827-860: for node j the rows i (ascending) with i -> j strong."""
    n = len(rp) - 1
    lists = [[] for _ in range(n)]
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            if s[k] and cv[k] != i:
                lists[cv[k]].append(i)
    return lists


def rs_first_pass(rp, cv, s, use_zpt, classical=False):
    """src/coarsening.jl:406-567: greedy Ruge-Stueben first pass over bucket
    lists keyed by lambda (head insertion; the head of the highest non-empty
    bucket is taken).  States: 1 C, -1 F, -2 tentative F (use_zpt).
    classical: coarsen_rs's own first pass (src/coarsening.jl:671-805): no
    zero-measure step and no marking when a decrement reaches zero."""
    n = len(rp) - 1
    f_pnt = -2 if use_zpt else -1
    lam = [int(c) for c in strong_transpose_count(rp, cv, s)]
    st_adj = _strong_transpose_adj(rp, cv, s)
    cf = [0] * n
    strong_rows = [[cv[k] for k in range(rp[i], rp[i + 1]) if s[k] and cv[k] != i] for i in range(n)]
    for i in range(n):
        if not strong_rows[i] and lam[i] == 0:
            cf[i] = 1
    for i in range(n):
        if classical or cf[i] != 0:
            continue
        if lam[i] == 0:
            cf[i] = f_pnt
            for j in strong_rows[i]:
                if cf[j] == 0:
                    lam[j] += 1
    max_lam = max([lam[i] for i in range(n) if cf[i] == 0 or classical], default=0)
    head = [-1] * max(max_lam + 1, 1)
    nxt = [-1] * n
    prv = [-1] * n
    for i in range(n):
        if cf[i] != 0:
            continue
        k = lam[i]
        h = head[k]
        head[k] = i
        nxt[i] = h
        prv[i] = -1
        if h != -1:
            prv[h] = i

    def remove(i):
        p, x = prv[i], nxt[i]
        if p != -1:
            nxt[p] = x
        else:
            head[lam[i]] = x
        if x != -1:
            prv[x] = p
        nxt[i] = prv[i] = -1

    def update(i, v):
        remove(i)
        lam[i] = v
        while v >= len(head):
            head.append(-1)
        h = head[v]
        head[v] = i
        nxt[i] = h
        prv[i] = -1
        if h != -1:
            prv[h] = i

    top = max_lam

    def raise_around(j):
        nonlocal top
        for k in strong_rows[j]:
            if cf[k] == 0:
                v = lam[k] + 1
                update(k, v)
                top = max(top, v)

    while True:
        while top >= 0 and head[top] == -1:
            top -= 1
        if top < 0:
            break
        best = head[top]
        remove(best)
        cf[best] = 1
        for j in st_adj[best]:
            if cf[j] != 0:
                continue
            remove(j)
            cf[j] = -1
            raise_around(j)
        for j in strong_rows[best]:
            if cf[j] != 0:
                continue
            v = max(0, lam[j] - 1)
            update(j, v)
            if v == 0 and not classical:
                remove(j)
                cf[j] = f_pnt
                raise_around(j)
    return np.array(cf, dtype=np.int64)


def pmis_on_undecided(cf, rp, cv, s, measure):
    """src/coarsening.jl:575-624, in place in row order (0 -> 1 / -1; leftovers -1)."""
    n = len(rp) - 1
    for _ in range(n + 1):
        all_decided = True
        for i in range(n):
            if cf[i] != 0:
                continue
            all_decided = False
            if measure[i] < 1.0:
                cf[i] = -1
                continue
            is_max = True
            for k in range(rp[i], rp[i + 1]):
                j = cv[k]
                if j != i and s[k] and cf[j] != -1 and measure[j] > measure[i]:
                    is_max = False
                    break
            if is_max:
                cf[i] = 1
        for i in range(n):
            if cf[i] != 0:
                continue
            for k in range(rp[i], rp[i + 1]):
                if s[k] and cf[cv[k]] == 1:
                    cf[i] = -1
                    break
        if all_decided:
            break
    cf[cf == 0] = -1
    return cf


def coarsen_hmis(rp, cv, nz, theta, seed):
    """src/coarsening.jl:330-394: RS first pass with tentative F points, then
    n draws, tentative F points with measure >= 1 or a strong connection
    reopened, PMIS on the undecided nodes."""
    n = len(rp) - 1
    if n <= 1:
        return np.ones(n, dtype=np.int64)
    s = strength_absolute(rp, cv, nz, theta)
    cf = rs_first_pass(rp, cv, s, True)
    measure = strong_transpose_count(rp, cv, s).astype(np.float64) + splitmix_uniforms(seed, n)
    for i in range(n):
        if cf[i] == -2:
            has = any(cv[k] != i and s[k] for k in range(rp[i], rp[i + 1]))
            cf[i] = 0 if (measure[i] >= 1.0 or has) else -1
    return pmis_on_undecided(cf, rp, cv, s, measure)


def ensure_fine_have_coarse_neighbor(cf, rp, cv, s):
    """src/coarsening.jl:291-313: F points without a strong C neighbour become
    C, sweeping in row order until nothing changes."""
    n = len(rp) - 1
    changed = True
    while changed:
        changed = False
        for i in range(n):
            if cf[i] != -1:
                continue
            if not any(cv[k] != i and s[k] and cf[cv[k]] == 1 for k in range(rp[i], rp[i + 1])):
                cf[i] = 1
                changed = True
    return cf


def coarsen_aggressive_cf(rp, cv, nz, theta, hmis_base, seed):
    """src/coarsening.jl:990-1116 (max_row_sum = 1: no weakening).  The first
    pass draws n values, the second pass one per C point, from the same
    stream (the reference's one generator)."""
    n = len(rp) - 1
    if n <= 1:
        return np.ones(n, dtype=np.int64)
    cf1 = (coarsen_hmis if hmis_base else coarsen_pmis)(rp, cv, nz, theta, seed)
    s = strength_absolute(rp, cv, nz, theta)
    c_idx = [i for i in range(n) if cf1[i] == 1]
    c_loc = {i: a for a, i in enumerate(c_idx)}
    m = len(c_idx)
    nbrs = [set() for _ in range(m)]
    for i in range(n):
        ci = [c_loc[cv[k]] for k in range(rp[i], rp[i + 1]) if cv[k] != i and s[k] and cf1[cv[k]] == 1]
        for a in ci:
            for b in ci:
                if a != b:
                    nbrs[a].add(b)
            if cf1[i] == 1:
                nbrs[c_loc[i]].add(a)
                nbrs[a].add(c_loc[i])
    draws = splitmix_uniforms(seed, n + m)[n:]  # after the first pass's n draws
    measure = np.array([len(nbrs[a]) for a in range(m)], dtype=np.float64) + draws
    cf2 = [0] * m
    for _ in range(m + 1):
        all_decided = True
        for a in range(m):
            if cf2[a] != 0:
                continue
            all_decided = False
            if all(not (cf2[b] != -1 and measure[b] > measure[a]) for b in nbrs[a]):
                cf2[a] = 1
        for a in range(m):
            if cf2[a] != 0:
                continue
            if any(cf2[b] == 1 for b in nbrs[a]):
                cf2[a] = -1
        if all_decided:
            break
    cf = cf1.copy()
    for a in range(m):
        if cf2[a] == -1:
            cf[c_idx[a]] = -1
    return ensure_fine_have_coarse_neighbor(cf, rp, cv, s)


_EPS = 2.220446049250313e-16


def interp_ext_i(rp, cv, nz, cf, theta, trunc=0.0, max_elements=0, norm_p=1, rescale=False):
    """src/prolongation.jl:467-690.  Returns P as (rp, cv, nz), rows sorted by
    column (src/prolongation.jl:713-734).  cf: 1 C, -1 F."""
    n = len(rp) - 1
    s = strength_absolute(rp, cv, nz, theta)
    cmap = np.cumsum(cf == 1) - 1
    strong = [[cv[k] for k in range(rp[i], rp[i + 1]) if cv[k] != i and s[k]] for i in range(n)]
    rows = []
    for i in range(n):
        if cf[i] == 1:
            rows.append([(int(cmap[i]), 1.0)])
            continue
        marker = {}  # j -> position in C-hat, or "F" for a strong F neighbour of i
        chat = []
        for j in strong[i]:
            if cf[j] == 1:
                if marker.get(j, -1) == -1:
                    marker[j] = len(chat)
                    chat.append(j)
            elif cf[j] == -1:
                marker[j] = "F"
                for k in strong[j]:
                    if cf[k] == 1 and marker.get(k, -1) == -1:
                        marker[k] = len(chat)
                        chat.append(k)
        if not chat:  # nearest coarse point by |a_ij|
            best, bv = 0, 0.0
            for k in range(rp[i], rp[i + 1]):
                j = cv[k]
                if j != i and cf[j] == 1 and abs(nz[k]) > bv:
                    bv, best = abs(nz[k]), int(cmap[j])
            rows.append([(best, 1.0)])
            continue
        pd = [0.0] * len(chat)
        diag = 0.0
        for k in range(rp[i], rp[i + 1]):
            j, a = cv[k], nz[k]
            if j == i:
                diag += a
                continue
            mk = marker.get(j, -1)
            if mk != -1 and mk != "F":
                pd[mk] += a
            elif mk == "F":
                dj = 0.0
                for k3 in range(rp[j], rp[j + 1]):
                    if cv[k3] == j:
                        dj = nz[k3]
                        break
                sgn = -1 if dj < 0 else 1
                sm = 0.0
                for k2 in range(rp[j], rp[j + 1]):
                    m = cv[k2]
                    if m == j:
                        continue
                    if sgn * nz[k2] < 0:
                        mm = marker.get(m, -1)
                        if (mm != -1 and mm != "F") or m == i:
                            sm += nz[k2]
                if abs(sm) > _EPS:
                    dist = a / sm
                    for k2 in range(rp[j], rp[j + 1]):
                        m = cv[k2]
                        if m == j:
                            continue
                        if sgn * nz[k2] < 0:
                            mm = marker.get(m, -1)
                            if mm != -1 and mm != "F":
                                pd[mm] += dist * nz[k2]
                            elif m == i:
                                diag += dist * nz[k2]
                else:
                    diag += a
            else:
                diag += a
        if abs(diag) > _EPS:
            pd = [v / -diag for v in pd]
        keep = list(range(len(chat)))
        if trunc > 0:
            mx = max(abs(v) ** norm_p for v in pd)
            thr = trunc * mx
            keep = [t for t in keep if abs(pd[t]) ** norm_p >= thr]
        if max_elements > 0 and len(keep) > max_elements:
            keep = sorted(keep, key=lambda t: abs(pd[t]), reverse=True)[:max_elements]
        scale = 1.0
        if rescale and len(keep) < len(chat):
            kept = set(keep)
            removed = 0.0
            for t in range(len(chat)):
                if t not in kept:
                    removed += pd[t]
            den = 1.0 - removed
            if abs(den) > 1e-12:
                scale = 1.0 / den
        rows.append(sorted((int(cmap[chat[t]]), pd[t] * scale) for t in keep))
    prp = np.zeros(n + 1, dtype=np.int64)
    for i, r in enumerate(rows):
        prp[i + 1] = prp[i] + len(r)
    pcv = np.array([c for r in rows for c, _ in r], dtype=np.int64)
    pnz = np.array([v for r in rows for _, v in r], dtype=np.float64)
    return prp, pcv, pnz


def filter_prolongation(prp, pcv, pnz, tol):
    """src/prolongation.jl:134-182: drop |p| < tol * max |p| per row (the
    largest kept if all would go); no rescaling."""
    n = len(prp) - 1
    orp, ocv, onz = [0], [], []
    for i in range(n):
        a, b = prp[i], prp[i + 1]
        if a == b:
            orp.append(len(ocv))
            continue
        thr = tol * max(abs(pnz[k]) for k in range(a, b))
        cnt = 0
        for k in range(a, b):
            if abs(pnz[k]) >= thr:
                ocv.append(pcv[k])
                onz.append(pnz[k])
                cnt += 1
        if cnt == 0:
            best, bv = a, 0.0
            for k in range(a, b):
                if abs(pnz[k]) > bv:
                    bv, best = abs(pnz[k]), k
            ocv.append(pcv[best])
            onz.append(pnz[best])
        orp.append(len(ocv))
    return np.array(orp, dtype=np.int64), np.array(ocv, dtype=np.int64), np.array(onz, dtype=np.float64)


def _sign(x):
    return 1 if x > 0 else (-1 if x < 0 else 0)


def _rows_to_csr(rows):
    n = len(rows)
    prp = np.zeros(n + 1, dtype=np.int64)
    for i, r in enumerate(rows):
        prp[i + 1] = prp[i] + len(r)
    return (prp, np.array([c for r in rows for c, _ in r], dtype=np.int64),
            np.array([v for r in rows for _, v in r], dtype=np.float64))


def interp_direct(rp, cv, nz, cf, theta):
    """src/prolongation.jl:219-335: strong C neighbours of opposite sign to the
    diagonal, w = -a_ij / d_i with d_i = a_ii + every other off-diagonal."""
    n = len(rp) - 1
    s = strength_absolute(rp, cv, nz, theta)
    cmap = np.cumsum(cf == 1) - 1
    rows = []
    for i in range(n):
        if cf[i] == 1:
            rows.append([(int(cmap[i]), 1.0)])
            continue
        a_ii, dsign = 0.0, 0
        for k in range(rp[i], rp[i + 1]):
            if cv[k] == i:
                a_ii = nz[k]
                dsign = _sign(a_ii)
        sum_non_c = 0.0
        cols, vals = [], []
        for k in range(rp[i], rp[i + 1]):
            j = cv[k]
            if j == i:
                continue
            if s[k] and cf[j] == 1 and (dsign == 0 or _sign(nz[k]) != dsign):
                cols.append(int(cmap[j]))
                vals.append(nz[k])
            else:
                sum_non_c += nz[k]
        d = a_ii + sum_non_c
        if not cols:
            best, bv = 0, 0.0
            for k in range(rp[i], rp[i + 1]):
                j = cv[k]
                if j != i and cf[j] == 1 and abs(nz[k]) > bv:
                    bv, best = abs(nz[k]), int(cmap[j])
            rows.append([(best, 1.0)])
            continue
        ok = abs(d) > _EPS * max(1.0, abs(d))
        rows.append([(c, (-v / d) if ok else 0.0) for c, v in zip(cols, vals)])
    return _rows_to_csr(rows)


def interp_standard(rp, cv, nz, cf, theta):
    """src/prolongation.jl:349-456: strong C neighbours directly, strong F
    neighbours k through their connections to C_i (and back to i) of sign
    opposite to a_kk; weak connections lumped into d_i."""
    n = len(rp) - 1
    s = strength_absolute(rp, cv, nz, theta)
    cmap = np.cumsum(cf == 1) - 1
    rows = []
    for i in range(n):
        if cf[i] == 1:
            rows.append([(int(cmap[i]), 1.0)])
            continue
        a_ii, sum_weak = 0.0, 0.0
        strong_c = {}
        strong_f = []
        for k in range(rp[i], rp[i + 1]):
            j = cv[k]
            if j == i:
                a_ii = nz[k]
            elif s[k] and cf[j] == 1:
                cm = int(cmap[j])
                strong_c[cm] = strong_c.get(cm, 0.0) + nz[k]
            elif s[k] and cf[j] == -1:
                strong_f.append((j, nz[k]))
            else:
                sum_weak += nz[k]
        d = a_ii + sum_weak
        contrib = dict(strong_c)
        for kk, a_ik in strong_f:
            diag_k = 0.0
            for k2 in range(rp[kk], rp[kk + 1]):
                if cv[k2] == kk:
                    diag_k = nz[k2]
                    break
            sgn = -1 if diag_k < 0 else 1
            sum_ck = 0.0
            cvals = {}
            for k2 in range(rp[kk], rp[kk + 1]):
                j2 = cv[k2]
                if j2 == kk:
                    continue
                a_kj = nz[k2]
                if cf[j2] == 1:
                    cm2 = int(cmap[j2])
                    if cm2 in strong_c and sgn * a_kj < 0:
                        cvals[cm2] = cvals.get(cm2, 0.0) + a_kj
                        sum_ck += a_kj
                if j2 == i and sgn * a_kj < 0:
                    sum_ck += a_kj
            if abs(sum_ck) > _EPS:
                dist = a_ik / sum_ck
                for cm2, a_kj in cvals.items():
                    contrib[cm2] = contrib.get(cm2, 0.0) + dist * a_kj
                for k2 in range(rp[kk], rp[kk + 1]):
                    if cv[k2] == i and sgn * nz[k2] < 0:
                        d += dist * nz[k2]
                        break
            else:
                d += a_ik
        if not contrib:
            best, bv = 0, 0.0
            for k in range(rp[i], rp[i + 1]):
                j = cv[k]
                if j != i and cf[j] == 1 and abs(nz[k]) > bv:
                    bv, best = abs(nz[k]), int(cmap[j])
            rows.append([(best, 1.0)])
            continue
        rows.append(sorted((cm, (-v / d) if abs(d) > _EPS else 0.0) for cm, v in contrib.items()))
    return _rows_to_csr(rows)


def coarsen_rs(rp, cv, nz, theta):
    """src/coarsening.jl:671-818: classical first pass, then every F point
    without a strong C neighbour promoted (:291-313)."""
    n = len(rp) - 1
    if n <= 1:
        return np.ones(n, dtype=np.int64)
    s = strength_absolute(rp, cv, nz, theta)
    cf = rs_first_pass(rp, cv, s, False, classical=True)
    return ensure_fine_have_coarse_neighbor(cf, rp, cv, s)


def smooth_prolongation(rp, cv, nz, prp, pcv, pnz, omega):
    """src/prolongation.jl:38-120: P = (I - omega D^-1 A) P_tent, entries with
    |p| <= eps dropped (the row's first tentative entry kept if all go)."""
    n = len(rp) - 1
    inv = np.zeros(n)
    for i in range(n):
        d = 0.0
        for k in range(rp[i], rp[i + 1]):
            if cv[k] == i:
                d = nz[k]
                break
        inv[i] = 1.0 / d if abs(d) > _EPS * max(1.0, abs(d)) else 0.0
    rows = []
    for i in range(n):
        ent = {}
        for q in range(prp[i], prp[i + 1]):
            ent[pcv[q]] = ent.get(pcv[q], 0.0) + pnz[q]
        factor = -omega * inv[i]
        for k in range(rp[i], rp[i + 1]):
            j = cv[k]
            w = factor * nz[k]
            for q in range(prp[j], prp[j + 1]):
                ent[pcv[q]] = ent.get(pcv[q], 0.0) + w * pnz[q]
        row = sorted((int(J), v) for J, v in ent.items() if abs(v) > _EPS)
        if not row and prp[i] < prp[i + 1]:
            row = [(int(pcv[prp[i]]), pnz[prp[i]])]
        rows.append(row)
    return _rows_to_csr(rows)


def coarsen_aggressive_agg(rp, cv, nz, theta, seed):
    """src/coarsening.jl:869-957, 1-based aggregate numbers inside (0 = none);
    returns 0-based aggregates and their count."""
    n = len(rp) - 1
    cf = coarsen_pmis(rp, cv, nz, theta, seed)
    s = strength_absolute(rp, cv, nz, theta)
    agg = [0] * n
    cnt = 0
    for i in range(n):
        if cf[i] == 1:
            cnt += 1
            agg[i] = cnt
    for i in range(n):
        if cf[i] == 1:
            continue
        best, bv = 0, 0.0
        for k in range(rp[i], rp[i + 1]):
            j = cv[k]
            if j != i and cf[j] == 1 and abs(nz[k]) > bv:
                bv, best = abs(nz[k]), agg[j]
        if best != 0:
            agg[i] = best
    parent = list(range(cnt + 1))  # 1-based, entry 0 unused

    def find(i):
        while parent[i] != i:
            parent[i] = parent[parent[i]]
            i = parent[i]
        return i

    def union(a, b):
        ra, rb = find(a), find(b)
        if ra != rb:
            parent[rb] = ra

    for i in range(n):
        if cf[i] != -1:
            continue
        roots = [find(agg[cv[k]]) for k in range(rp[i], rp[i + 1]) if s[k] and cf[cv[k]] == 1]
        for r in roots[1:]:
            union(roots[0], r)
    for i in range(1, cnt + 1):
        find(i)
    new_id = [0] * (cnt + 1)
    nc = 0
    for i in range(1, cnt + 1):
        r = find(i)
        if new_id[r] == 0:
            nc += 1
            new_id[r] = nc
        new_id[i] = new_id[r]
    for i in range(n):
        if agg[i] > 0:
            agg[i] = new_id[find(agg[i])]
    for i in range(n):
        if agg[i] == 0:
            best, bv = 0, 0.0
            for k in range(rp[i], rp[i + 1]):
                j = cv[k]
                if j != i and agg[j] != 0 and abs(nz[k]) > bv:
                    bv, best = abs(nz[k]), agg[j]
            if best != 0:
                agg[i] = best
            else:
                nc += 1
                agg[i] = nc
    return np.array(agg, dtype=np.int64) - 1, nc


def apply_max_row_sum(rp, cv, nz, threshold):
    """src/strength.jl:183-213: rows with |row sum| > |a_ii| * threshold
    (threshold < 1) lose their off-diagonal values (set to 0, pattern kept)."""
    n = len(rp) - 1
    out = np.array(nz, dtype=np.float64, copy=True)
    for i in range(n):
        d, rs = 0.0, 0.0
        for k in range(rp[i], rp[i + 1]):
            rs += nz[k]
            if cv[k] == i:
                d = nz[k]
        if abs(d) < _EPS:
            continue
        if abs(rs) > abs(d) * threshold and threshold < 1.0:
            for k in range(rp[i], rp[i + 1]):
                if cv[k] != i:
                    out[k] = 0.0
    return out
