"""Oracle: the whole-cell JD certificate restated in NumPy (test infrastructure only).

Follows jd_curl/three_d/interval.py:31-124 (binary64 intervals, one outward nextafter per elementary
operation), feasibility.py:188-217 (edge hull), :239-290 (jacobian_intervals, _sample_jacobians),
:305-372 (_classify, adaptive lowest-bound-first policy), :395-431 and :571-769 (certify_map with the
prescreen configuration) and compiled_enclosure.py:193-227 (hoisted edge-hull margins). Vectorised over
cells; the subdivision loop runs per cell in Python (only the few undecided cells reach it).

Pinned in tests/test_oracle_certificate.py by exact rational arithmetic (every interval contains the
exact value of its operation; the cell intervals contain the exact J and Bernstein coefficients;
certified cells have exact J > delta on a lattice, witnesses have exact J <= delta), closed forms
(identity and affine maps) and the T05 folding cell.
"""
import heapq
from dataclasses import dataclass

import numpy as np

from . import prescreen as opre

INF = np.inf


def outward(lo, hi):
    """(nextafter(lo, -inf), nextafter(hi, +inf)) and validity (finite, ordered)."""
    with np.errstate(all="ignore"):
        lo, hi = np.nextafter(lo, -INF), np.nextafter(hi, INF)
        ok = np.isfinite(lo) & np.isfinite(hi) & (lo <= hi)
    return lo, hi, ok


def iadd(a, b):
    with np.errstate(all="ignore"):
        return outward(a[0] + b[0], a[1] + b[1])


def isub(a, b):
    with np.errstate(all="ignore"):
        return outward(a[0] - b[1], a[1] - b[0])


def imul(a, b):
    with np.errstate(all="ignore"):
        x, y, z, w = a[0] * b[0], a[0] * b[1], a[1] * b[0], a[1] * b[1]
        return outward(np.minimum(np.minimum(np.minimum(x, y), z), w),
                       np.maximum(np.maximum(np.maximum(x, y), z), w))


class Chain:
    """Interval values (lo, hi) with the AND of every step's validity."""

    def __init__(self):
        self.ok = True

    def __call__(self, res):
        lo, hi, ok = res
        ok = np.asarray(ok)
        if ok.ndim > 1:
            ok = ok.reshape(len(ok), -1).all(axis=1)     # per cell
        self.ok = self.ok & ok
        return lo, hi


def idet(F, ch):
    """Left-associated cofactor expansion along row 0 of an interval matrix F[r][j] = (lo, hi)."""
    a = ch(isub(ch(imul(F[1][1], F[2][2])), ch(imul(F[1][2], F[2][1]))))
    b = ch(isub(ch(imul(F[1][2], F[2][0])), ch(imul(F[1][0], F[2][2]))))
    c = ch(isub(ch(imul(F[1][0], F[2][1])), ch(imul(F[1][1], F[2][0]))))
    return ch(iadd(ch(iadd(ch(imul(F[0][0], a)), ch(imul(F[0][1], b)))), ch(imul(F[0][2], c))))


def cell_vertices(u, grid_shape, cells=None):
    """(n, 8, 3) vertex displacements of the cells (vertex a = x + 2y + 4z)."""
    nz, ny, nx = grid_shape
    cz, cy, cx = nz - 1, ny - 1, nx - 1
    cells = np.arange(cz * cy * cx) if cells is None else np.asarray(cells)
    z0, y0, x0 = cells // (cx * cy), (cells // cx) % cy, cells % cx
    flat = np.asarray(u, dtype=np.float64).reshape(-1, 3)
    idx = np.stack([((z0 + (a >> 2)) * ny + y0 + ((a >> 1) & 1)) * nx + x0 + (a & 1) for a in range(8)], axis=-1)
    return flat[idx]


def edge_hull_margins(vertices, spacing, delta):
    """Per cell lo(J_hull - delta) and the invalid-step flag (compiled_enclosure.py:193-227)."""
    v = np.asarray(vertices, dtype=np.float64)
    ch = Chain()
    F = [[None] * 3 for _ in range(3)]
    for j in range(3):
        lows = [a for a in range(8) if not a & (1 << j)]
        for r in range(3):
            with np.errstate(all="ignore"):
                d = np.stack([v[:, a | (1 << j), r] - v[:, a, r] for a in lows], axis=-1)
            finite = np.all(np.isfinite(d), axis=-1)
            ch.ok = ch.ok & finite
            lower, upper = d[:, 0], d[:, 0]
            for k in range(1, 4):
                lower, upper = np.minimum(lower, d[:, k]), np.maximum(upper, d[:, k])
            e = ch(outward(lower, upper))
            with np.errstate(all="ignore"):
                e = ch(outward(e[0] / spacing[j], e[1] / spacing[j]))
            one = 1. if r == j else 0.
            F[r][j] = ch(iadd(e, (one, one)))
    J = idet(F, ch)
    with np.errstate(all="ignore"):
        m = ch(outward(J[0] - delta, J[1] - delta))
    ok = np.broadcast_to(ch.ok, m[0].shape)
    return m[0], ~ok


POINTS = np.array([(x, y, z) for z in (0., .5, 1.) for y in (0., .5, 1.) for x in (0., .5, 1.)])


def edge_weights(j):
    """[(a, b, weight(27))] of the four j-edges in vertex order (exact dyadic bilinear weights)."""
    out = []
    for a in range(8):
        if a & (1 << j):
            continue
        w = np.ones(27)
        for k in range(3):
            if k != j:
                w = w * (POINTS[:, k] if (a >> k) & 1 else 1 - POINTS[:, k])
        out.append((a, a | (1 << j), w))
    return out


def bernstein_arrays(vertices, spacing, delta):
    """(p, c, samples, ok): p = J - delta at the 27 points and c its Bernstein coefficients as (lo, hi) pairs of
    (n, 27) arrays, float64 samples (n, 27); ok per cell (feasibility.py:239-290, interval.py:106-113)."""
    v = np.asarray(vertices, dtype=np.float64)
    n = len(v)
    ch = Chain()
    F = [[None] * 3 for _ in range(3)]
    for j in range(3):
        h = spacing[j]
        acc = [(np.zeros((n, 27)), np.zeros((n, 27))) for _ in range(3)]
        for a, b, w in edge_weights(j):
            for r in range(3):
                with np.errstate(all="ignore"):
                    d = v[:, b, r] - v[:, a, r]
                s = ch(outward(d, d))
                with np.errstate(all="ignore"):
                    x1, x2 = s[0] / h, s[1] / h
                q = ch(outward(np.minimum(np.minimum(np.minimum(x1, x1), x2), x2),
                               np.maximum(np.maximum(np.maximum(x1, x1), x2), x2)))
                with np.errstate(all="ignore"):
                    t = ch(outward(q[0][:, None] * w[None, :], q[1][:, None] * w[None, :]))
                acc[r] = ch(iadd(acc[r], t))
        for r in range(3):
            one = 1. if r == j else 0.
            F[r][j] = ch(iadd(acc[r], (one, one)))
    J = idet(F, ch)
    p = ch(isub(J, (delta, delta)))
    lo, hi = p[0].copy(), p[1].copy()
    for st in (1, 3, 9):
        mids = [i for i in range(27) if (i // st) % 3 == 1]
        for m in mids:
            a0, a2 = m - st, m + st
            two = ch(imul((lo[:, m], hi[:, m]), (2., 2.)))
            s = ch(iadd((lo[:, a0], hi[:, a0]), (lo[:, a2], hi[:, a2])))
            half = ch(imul(s, (.5, .5)))
            lo[:, m], hi[:, m] = ch(isub(two, half))
    samples, sok = sample_jacobians(v, spacing)
    ok = np.broadcast_to(ch.ok, (n,)) & sok
    return p, (lo, hi), samples, ok


def sample_jacobians(v, spacing):
    """float64 J at the 27 points (feasibility.py:271-288): F = I + sum (diff/h) w in edge order, cofactor rows,
    det = ((0 + f0 a0) + f1 a1) + f2 a2; ok = every entry, cofactor and product finite."""
    n = len(v)
    F = np.zeros((n, 27, 3, 3))
    F[..., 0, 0] = F[..., 1, 1] = F[..., 2, 2] = 1.
    with np.errstate(all="ignore"):
        for j in range(3):
            for a, b, w in edge_weights(j):
                for r in range(3):
                    F[..., r, j] = F[..., r, j] + ((v[:, b, r] - v[:, a, r]) / spacing[j])[:, None] * w[None, :]
        f = F
        cof = [f[..., 1, 1] * f[..., 2, 2] - f[..., 1, 2] * f[..., 2, 1],
               f[..., 1, 2] * f[..., 2, 0] - f[..., 1, 0] * f[..., 2, 2],
               f[..., 1, 0] * f[..., 2, 1] - f[..., 1, 1] * f[..., 2, 0],
               f[..., 2, 1] * f[..., 0, 2] - f[..., 2, 2] * f[..., 0, 1],
               f[..., 2, 2] * f[..., 0, 0] - f[..., 2, 0] * f[..., 0, 2],
               f[..., 2, 0] * f[..., 0, 1] - f[..., 2, 1] * f[..., 0, 0],
               f[..., 0, 1] * f[..., 1, 2] - f[..., 0, 2] * f[..., 1, 1],
               f[..., 0, 2] * f[..., 1, 0] - f[..., 0, 0] * f[..., 1, 2],
               f[..., 0, 0] * f[..., 1, 1] - f[..., 0, 1] * f[..., 1, 0]]
        t = [f[..., 0, 0] * cof[0], f[..., 0, 1] * cof[1], f[..., 0, 2] * cof[2]]
        s = ((0. + t[0]) + t[1]) + t[2]
    fin = np.all(np.isfinite(F), axis=(-2, -1)) & np.all([np.isfinite(c) for c in cof], axis=0)
    fin &= np.all([np.isfinite(x) for x in t], axis=0) & np.isfinite(s)
    return s, fin.all(axis=1)


@dataclass
class CellResult:
    status: str
    reason: str
    margin: float | None
    lower: float | None
    boxes: int
    depth: int
    witness: tuple | None


def _split(lo, hi, axis):
    """Midpoint de Casteljau of (27,) Bernstein intervals along reference axis 0/1/2 (interval.py:116-124)."""
    st = (1, 3, 9)[axis]
    ch = Chain()
    L = (lo.copy(), hi.copy())
    R = (lo.copy(), hi.copy())
    for i0 in range(27):
        if (i0 // st) % 3:
            continue
        a, b, c = ((lo[i], hi[i]) for i in (i0, i0 + st, i0 + 2 * st))
        ab = ch(imul(ch(iadd(a, b)), (.5, .5)))
        bc = ch(imul(ch(iadd(b, c)), (.5, .5)))
        mid = ch(imul(ch(iadd(ab, bc)), (.5, .5)))
        for arr, vals in ((L, (a, ab, mid)), (R, (mid, bc, c))):
            for k, val in enumerate(vals):
                arr[0][i0 + k * st], arr[1][i0 + k * st] = val
    return L, R, bool(np.all(ch.ok))


def classify(p, c, delta, max_depth=48, max_boxes=255):
    """_classify with the adaptive lowest-bound-first policy for one cell; p, c are (lo(27,), hi(27,))."""
    root_lower = float(np.min(c[0]))

    def result(status, margin, boxes, depth, witness, reason):
        lower = margin
        if delta != 0:
            lo, hi, ok = iadd((margin, margin), (delta, delta))
            if not ok:
                return CellResult("uncertain", "arithmetic_uncertain", None, None, boxes, depth, None)
            lower = float(lo)
        return CellResult(status, reason, margin, lower, boxes, depth, witness)

    for idx in range(27):
        if p[1][idx] <= 0:
            iz, iy, ix = idx // 9, (idx // 3) % 3, idx % 3
            return result("infeasible", root_lower, 1, 0, (ix / 2, iy / 2, iz / 2), "sample_witness")
    heap = [(root_lower, 0, (c[0].copy(), c[1].copy()), np.zeros(3), np.ones(3), 0)]
    serial = boxes = depth_used = 0
    uncertain = False
    lowers = []
    while heap:
        if boxes >= max_boxes:
            return result("uncertain", root_lower, boxes, depth_used, None, "box_budget")
        _, _, cur, low, high, depth = heapq.heappop(heap)
        boxes += 1
        depth_used = max(depth_used, depth)
        minimum = float(np.min(cur[0]))
        if minimum > 0:
            lowers.append(minimum)
            continue
        for kz in (0, 1):
            for ky in (0, 1):
                for kx in (0, 1):
                    if cur[1][9 * 2 * kz + 3 * 2 * ky + 2 * kx] <= 0:
                        w = tuple(float(hv if k else lv) for k, hv, lv in zip((kx, ky, kz), high, low))
                        return result("infeasible", root_lower, boxes, depth_used, w, "subdivision_witness")
        if depth >= max_depth:
            uncertain = True
            continue
        axis = int(np.argmax(high - low))
        left, right, ok = _split(cur[0], cur[1], axis)
        if not ok:
            return result("uncertain", root_lower, boxes, depth_used, None, "arithmetic_uncertain")
        mid = (low[axis] + high[axis]) * .5
        lh, rl = high.copy(), low.copy()
        lh[axis] = rl[axis] = mid
        for child, lo_, hi_ in ((left, low, lh), (right, rl, high)):
            serial += 1
            heapq.heappush(heap, (float(np.min(child[0])), serial, child, lo_, hi_, depth + 1))
    if uncertain:
        return result("uncertain", root_lower, boxes, depth_used, None, "depth_budget")
    return result("certified_feasible", min(lowers), boxes, depth_used, None, "positive_bernstein_cover")


def certify_map(u, shape, spacing, axes, delta, D=None, tile=8192, max_depth=48, max_boxes=255):
    """certify_map with the prescreen configuration (feasibility.py:395-431, 571-769) as a dict of the
    FeasibilityResult fields (None as None)."""
    u = np.asarray(u, dtype=np.float64)
    nz, ny, nx = shape
    ncell = (nz - 1) * (ny - 1) * (nx - 1)
    base = dict(status="uncertain", min_jd_lower_bound=None, min_margin_lower_bound=None, sampled_min_jd=None,
                n_cells_checked=0, n_uncertain=ncell, n_infeasible=0, n_boxes=0, max_depth=0, witness_cell=None,
                witness_reference_point=None, n_fast_cells=0, n_fallback_cells=0, n_prescreen_cells=0)
    if not np.all(np.isfinite(u)):
        return dict(base, reason="NONFINITE_STATE")
    if (np.any(u[[0, -1]] != 0) or np.any(u[:, [0, -1]] != 0) or np.any(u[:, :, [0, -1]] != 0)):
        return dict(base, status="infeasible", reason="NONZERO_BOUNDARY")
    bad = np.zeros(shape, dtype=bool)
    with np.errstate(all="ignore"):
        for j, coords in enumerate(axes):
            sh = [1, 1, 1]
            sh[2 - j] = len(coords)
            val = u[..., j] + np.asarray(coords).reshape(sh)
            bad |= ~np.isfinite(val) | (val < coords[0]) | (val > coords[-1])
            inner = val[1:-1, 1:-1, 1:-1]
            bad[1:-1, 1:-1, 1:-1] |= (inner <= coords[0]) | (inner >= coords[-1])
    if np.any(bad):
        return dict(base, status="infeasible", reason="INVALID_DOMAIN")
    verts = cell_vertices(u, shape)
    margin, flag = edge_hull_margins(verts, spacing, delta)
    margins = []
    fast = prescreened = unknown = bad_n = boxes = depth = 0
    arith = False
    sampled = INF
    witness_cell = witness = None
    jobs = []
    for s in range(0, ncell, tile):
        e = min(ncell, s + tile)
        proven = np.zeros(e - s, dtype=bool) if flag[s:e].any() else margin[s:e] > 0
        if proven.any():
            margins.append(float(margin[s:e][proven].min()))
        fast += int(proven.sum())
        rest = s + np.flatnonzero(~proven)
        if len(rest):
            jobs.append(rest)
    boxes = fast
    if D is not None and jobs:
        cells = np.concatenate(jobs)
        low = opre.lower_bounds_c(u, _Grid(shape, spacing), D, cells)
        with np.errstate(invalid="ignore"):
            mp = low if delta == 0 else np.nextafter(low - delta, -INF)
            ok = mp > 0
        kept, start = [], 0
        for rest in jobs:
            stop = start + len(rest)
            here = ok[start:stop]
            if here.any():
                margins.append(float(mp[start:stop][here].min()))
                prescreened += int(here.sum())
                boxes += int(here.sum())
            if (~here).any():
                kept.append(rest[~here])
            start = stop
        jobs = kept
    fallback = sum(len(j) for j in jobs)
    for cells in jobs:
        p, c, samples, ok = bernstein_arrays(verts[cells], spacing, delta)
        if not ok.all():
            arith = True
            unknown += len(cells)
            continue
        sampled = min(sampled, float(samples.min()))
        minima = c[0].min(axis=1)
        safe = minima > 0
        boxes += len(cells)
        if safe.any():
            margins.append(float(minima[safe].min()))
        for k in np.flatnonzero(~safe):
            r = classify((p[0][k], p[1][k]), (c[0][k], c[1][k]), delta, max_depth, max_boxes)
            if r.margin is not None:
                margins.append(r.margin)
            arith = arith or r.reason == "arithmetic_uncertain"
            boxes += r.boxes - 1
            depth = max(depth, r.depth)
            unknown += r.status == "uncertain"
            bad_n += r.status == "infeasible"
            if r.status == "infeasible" and witness_cell is None:
                witness_cell, witness = int(cells[k]), r.witness
    status = "infeasible" if bad_n else ("uncertain" if unknown else "certified_feasible")
    lowest = min(margins) if margins and not arith else None
    lower = None
    if lowest is not None:
        lower = lowest if delta == 0 else float(iadd((lowest, lowest), (delta, delta))[0])
    reason = ("JACOBIAN_VIOLATION" if bad_n else "arithmetic_uncertain" if arith else
              "subdivision_budget" if unknown else "positive_bernstein_cover")
    if fast or prescreened:
        sampled = INF
        if status == "certified_feasible":
            reason = "positive_enclosure_cover"
    return dict(status=status, reason=reason, min_jd_lower_bound=lower, min_margin_lower_bound=lowest,
                sampled_min_jd=float(sampled) if np.isfinite(sampled) else None, n_cells_checked=ncell,
                n_uncertain=int(unknown), n_infeasible=int(bad_n), n_boxes=int(boxes), max_depth=int(depth),
                witness_cell=witness_cell, witness_reference_point=witness, n_fast_cells=int(fast),
                n_fallback_cells=int(fallback), n_prescreen_cells=int(prescreened))


@dataclass(frozen=True)
class _Grid:
    shape: tuple
    spacing: tuple
