"""Ordered FTLE ridges, restated in Python -- TEST INFRASTRUCTURE ONLY.

Follows the reference's extraction/ridges.py:
  ridge_connect_fields  _ftle_ridge_pts_connect   ridges.py:232-319
  _step                 _link_points_stepper      ridges.py:321-415
  linked_ridge_pts      _linked_ridge_pts         ridges.py:418-603
  _connect              _connect_endpoints        ridges.py:646-717
  ordered_ridges        ftle_ordered_ridges       ridges.py:720-1054

with two stated choices where the reference leaves behaviour to numpy or
to an exception: ties of the second directional derivative are ordered by
grid index (numpy's quicksort leaves them unordered), and the joining
step's acos arguments are clamped to [-1, 1] (math.acos raises on a
rounding excursion).  Pure-Python loops: small grids only.
"""
import math

import numpy as np


def ridge_connect_fields(f, eigvec_max, x, y, sdd_thresh=0.0, percentile=0):
    """Dense (nx*ny) ridge points (-1 elsewhere), eigenvectors and second
    directional derivatives (0 elsewhere), and h = min(dx, dy)."""
    f = np.asarray(f, dtype=np.float64)
    nx, ny = f.shape
    dx, dy = x[1] - x[0], y[1] - y[0]
    f_min = 0.0 if percentile == 0 else np.percentile(f, percentile)
    pts = -np.ones((nx * ny, 2))
    sdd = np.zeros(nx * ny)
    for i in range(2, nx - 2):
        for j in range(2, ny - 2):
            f0 = f[i, j]
            if not f0 > f_min:
                continue
            fx = (f[i + 1, j] - f[i - 1, j]) / (2 * dx)
            fy = (f[i, j + 1] - f[i, j - 1]) / (2 * dy)
            fxx = (f[i + 1, j] - 2 * f[i, j] + f[i - 1, j]) / (dx ** 2)
            fyy = (f[i, j + 1] - 2 * f[i, j] + f[i, j - 1]) / (dy ** 2)
            fxy = (f[i + 1, j + 1] - f[i + 1, j - 1] - f[i - 1, j + 1] + f[i - 1, j - 1]) / (
                4 * dx * dy)
            ex, ey = eigvec_max[i, j]
            c2 = ex * (fxx * ex + fxy * ey) + ey * (fxy * ex + fyy * ey)
            if c2 < -sdd_thresh:
                t = -(fx * ex + fy * ey) / c2
                if abs(t * ex) <= dx / 2 and abs(t * ey) <= dy / 2:
                    k = i * ny + j
                    pts[k] = (x[i] + t * ex, y[j] + t * ey)
                    sdd[k] = c2
    return pts, np.asarray(eigvec_max, dtype=np.float64).reshape(-1, 2).copy(), sdd, min(dx, dy)


def _acos_nan(v):
    # numba's math.acos returns NaN outside [-1, 1]
    return math.acos(v) if -1.0 <= v <= 1.0 else float("nan")


def _step(pt0, nvec0, ind0, pts, vecs, sdd, ridge_id, offsets, ridge_num, a, c):
    angle_ind = math.floor((math.atan2(nvec0[0], -nvec0[1]) % (2 * math.pi)) * 4 / math.pi
                           + 0.5) % 8
    cand = offsets[angle_ind:angle_ind + 3]
    metric = np.full(3, 10000.0)
    nvecs = np.zeros((3, 2))
    found = False
    for i in range(3):
        k = ind0 + cand[i]
        if sdd[k] < 0:
            d = ((pts[k, 0] - pt0[0]) ** 2 + (pts[k, 1] - pt0[1]) ** 2) ** 0.5
            n1 = vecs[k]
            dot = nvec0[0] * n1[0] + nvec0[1] * n1[1]
            if dot < 0:
                nvecs[i] = -n1
                dot = -dot
            else:
                nvecs[i] = n1
            metric[i] = a * d + c * _acos_nan(dot)
            found = True
    if not found:
        return None
    b = int(np.argmin(metric))
    k = ind0 + cand[b]
    if ridge_id[k] + 0.5 < 0:
        ridge_id[k] = ridge_num
        return pts[k].copy(), nvecs[b].copy(), k
    return None


def linked_ridge_pts(pts, vecs, sdd, nx, ny, h, sdd_thresh=0.0, c=1.0):
    """(ordered points, ridge_len (end, len), endpoints (x, y, tag),
    endpoint tangents) as _linked_ridge_pts returns them."""
    a = 1 / h
    order = np.argsort(sdd, kind="stable")
    offsets = np.array([di * ny + dj for di, dj in
                        [(1, -1), (1, 0), (1, 1), (0, 1), (-1, 1), (-1, 0), (-1, -1), (0, -1),
                         (1, -1), (1, 0)]])
    ridge_id = -np.ones(nx * ny)
    linked, ridge_len, endpoints, eptan = [], [], [], []
    num = 0
    for si in order:
        if not ridge_id[si] + 0.5 < 0:
            continue
        if not sdd[si] < -sdd_thresh:
            break
        ridge_id[si] = num
        start = len(linked)
        walk = [pts[si].copy()]
        p, n, k = pts[si].copy(), vecs[si].copy(), si
        for _ in range(10000):
            r = _step(p, n, k, pts, vecs, sdd, ridge_id, offsets, num, a, c)
            if r is None:
                break
            p, n, k = r
            walk.append(p)
        ep0 = walk[-1]
        walk.reverse()
        p, n, k = pts[si].copy(), -vecs[si], si
        ep1 = p
        for _ in range(10000):
            r = _step(p, n, k, pts, vecs, sdd, ridge_id, offsets, num, a, c)
            if r is None:
                break
            p, n, k = r
            ep1 = p
            walk.append(p)
        if len(walk) == 1:
            continue
        linked.extend(walk)
        t0 = walk[1] - ep0
        t0 = t0 / (t0[0] ** 2 + t0[1] ** 2) ** 0.5
        t1 = ep1 - walk[-2]
        t1 = t1 / (t1[0] ** 2 + t1[1] ** 2) ** 0.5
        ridge_len.append((len(linked), len(linked) - start))
        endpoints += [(ep0[0], ep0[1], num), (ep1[0], ep1[1], -(num + 0.1))]
        eptan += [t0, t1]
        num += 1
    return (np.array(linked).reshape(-1, 2), np.array(ridge_len, dtype=np.int64).reshape(-1, 2),
            np.array(endpoints).reshape(-1, 3), np.array(eptan).reshape(-1, 2))


def _acos_clamped(v):
    return math.acos(min(1.0, max(-1.0, v)))


def _distances(pt, arr, tol):
    tan = arr - pt
    dist = np.sqrt(tan[:, 0] ** 2 + tan[:, 1] ** 2)
    within = dist < tol
    tan[within] /= dist[within, None]
    return dist, tan, within


def _connect(endpoint, tan_vec, rem, rem_tan, ang, tol):
    dist, tan, within = _distances(endpoint[:2], rem[:, :2], tol)
    if not within.any():
        return rem, rem_tan, None, None, tan_vec
    sign = math.copysign(1, endpoint[2])
    for _ in range(int(within.sum())):
        k = int(np.argmin(dist))
        nsign = math.copysign(1, rem[k, 2])
        if (_acos_clamped(-sign * (tan[k] @ tan_vec)) < ang
                and _acos_clamped(nsign * (rem_tan[k] @ tan[k])) < ang):
            tag = rem[k, 2]
            p = k + int(math.copysign(1, tag))
            new_ep, new_tan = rem[p].copy(), rem_tan[p].copy()
            keep = np.ones(len(rem), bool)
            keep[[k, p]] = False
            return rem[keep], rem_tan[keep], tag, new_ep, new_tan
        dist[k] = 10 * tol
    return rem, rem_tan, None, None, tan_vec


def _chain(ep, tan, rem, rem_tan, ang, tol, nridges):
    tags = [-(ep[2] + 0.1)]
    for _ in range(nridges):
        rem, rem_tan, tag, ep, tan = _connect(ep, tan, rem, rem_tan, ang, tol)
        if ep is None:
            break
        tags.append(tag)
    return tags, rem, rem_tan


def _piece(linked, ridge_len, tag):
    r = int(round(abs(tag)))
    end, ln = ridge_len[r]
    seg = linked[end - ln:end]
    return seg[::-1] if tag + 0.01 < 0 else seg


def ordered_ridges(f, eigvec_max, x, y, dist_tol, ep_tan_ang=math.pi / 4, min_ridge_pts=5,
                   sdd_thresh=0.0, percentile=0, c=1.0, fields=None):
    """ftle_ordered_ridges restated; fields = ridge_connect_fields(...) may be
    passed in to link given dense buffers."""
    f = np.asarray(f, dtype=np.float64)
    nx, ny = f.shape
    if fields is None:
        fields = ridge_connect_fields(f, eigvec_max, x, y, sdd_thresh, percentile)
    pts, vecs, sdd, h = fields
    linked, ridge_len, endpoints, eptan = linked_ridge_pts(pts, vecs, sdd, nx, ny, h, sdd_thresh,
                                                           c)
    rem, rem_tan = endpoints.copy(), eptan.copy()
    nridges = len(ridge_len)
    out = []
    while len(rem):
        cur, cur_tan = rem[:2].copy(), rem_tan[:2].copy()
        rem, rem_tan = rem[2:], rem_tan[2:]
        d0, _, w0 = _distances(cur[0, :2], rem[:, :2], dist_tol)
        d1, _, w1 = _distances(cur[1, :2], rem[:, :2], dist_tol)
        case = int(w0.any()) + 2 * int(w1.any())
        single = None
        if case == 3:
            side = int(np.argmin([d0.min(), d1.min()]))
            t0, rem, rem_tan = _chain(cur[side], cur_tan[side], rem, rem_tan, ep_tan_ang,
                                      dist_tol, nridges)
            other = side + int(math.copysign(1, cur[side, 2]))
            t1, rem, rem_tan = _chain(cur[other], cur_tan[other], rem, rem_tan, ep_tan_ang,
                                      dist_tol, nridges)
            if len(t0) > 1 and len(t1) > 1:
                tags = [-(t + 0.1) for t in t0[::-1]] + t1[1:]
            elif len(t0) > 1:
                tags = t0
            else:
                tags = t1
            if len(t0) + len(t1) <= 2:
                single, check = t1[0], True
        elif case in (1, 2):
            side = 1 if case == 2 else 0
            tags, rem, rem_tan = _chain(cur[side], cur_tan[side], rem, rem_tan, ep_tan_ang,
                                        dist_tol, nridges)
            if len(tags) == 1:
                single, check = tags[0], case == 2
        else:
            single, check = cur[0, 2], True
        if single is not None:
            r = int(round(abs(single)))
            end, ln = ridge_len[r]
            if check and ln < min_ridge_pts:
                continue
            out.append(linked[end - ln:end].copy())
            continue
        pieces = [_piece(linked, ridge_len, t) for t in tags]
        if sum(len(p) for p in pieces) < min_ridge_pts:
            continue
        out.append(np.concatenate(pieces))
    return out
