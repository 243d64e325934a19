// Ordered FTLE ridges (ftle_ordered_ridges, extraction/ridges.py:720-1054 of
// the reference): ridge points linked into ordered curves by a greedy walk
// from the strongest points (_linked_ridge_pts, ridges.py:418-603, with the
// neighbour step _link_points_stepper, ridges.py:321-415), then curves whose
// endpoints are close and aligned joined (_connect_endpoints,
// ridges.py:646-717).  The walk is inherently serial (each step claims the
// point the next one may not take), so this is host code run on the dense
// per-point buffers the GPU ridge predicate produced.
//
// Conventions kept from the reference:
//   * points are visited in ascending second directional derivative (most
//     negative first); ties, which numpy's quicksort leaves unordered, go
//     by grid index;
//   * a step looks at the 3 neighbours around the ridge tangent, scores them
//     by a*distance + c*angle between normals (a = 1/min(dx, dy)), takes the
//     best (numpy argmin: a NaN score, from acos of a dot product rounded
//     above 1, wins) and stops if that point already belongs to a ridge;
//   * a ridge of one point is dropped; endpoints carry tags r (start) and
//     -(r + 0.1) (end); joined pieces are flipped where their tag is
//     negative;
//   * the angle tests of the joining step use acos of dot products of unit
//     vectors, clamped to [-1, 1] here (the reference's math.acos raises on
//     a rounding excursion beyond 1);
//   * a curve that only its start endpoint could extend, and which found no
//     partner, is kept whatever its length (the reference's branch for that
//     case has no min_ridge_pts test).
//
// Two ways to the same decisions at lower cost: only the points below the
// second-derivative threshold are sorted (the walk stops at the first other
// one), and the joining step finds the endpoints within dist_tol through a
// uniform grid instead of scanning all remaining endpoints per query.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

namespace nbx {
namespace ordered {

constexpr double kPi = 3.141592653589793;

// Dense per-grid-point buffers of the ridge predicate (nx*ny entries each).
struct RidgeField {
    int nx = 0, ny = 0;
    double h = 0;                // min(dx, dy)
    std::vector<double> px, py;  // ridge point (valid where sdd < 0)
    std::vector<double> vx, vy;  // eigvec_max at the grid point
    std::vector<double> sdd;     // second directional derivative, 0 if not a ridge point
};

struct Linked {
    std::vector<double> pts;       // ordered points, x/y interleaved
    std::vector<int64_t> end, len; // per ridge: end index (exclusive) and length
    std::vector<double> ep;        // 2 per ridge: x, y, tag
    std::vector<double> eptan;     // 2 per ridge: tx, ty
};

inline double py_mod(double a, double m) {
    double r = std::fmod(a, m);
    if (r != 0.0 && ((r < 0) != (m < 0))) r += m;
    return r;
}

// numpy.argmin over n values: the first NaN if any, else the first minimum
inline int argmin_np(const double* v, int n) {
    int best = 0;
    for (int i = 0; i < n; ++i) {
        if (std::isnan(v[i])) return i;
        if (v[i] < v[best]) best = i;
    }
    return best;
}

// One step of the walk (_link_points_stepper).  Returns false when the walk
// ends; otherwise moves (p, n, ind) to the claimed neighbour.
inline bool step(const RidgeField& F, std::vector<double>& tag, double& p0x, double& p0y,
                 double& n0x, double& n0y, int64_t& ind0, const int64_t off[10], double ridge_num,
                 double a, double c) {
    const double ang = py_mod(std::atan2(n0x, -n0y), 2 * kPi);
    const int ai = static_cast<int>(std::floor(ang * 4 / kPi + 0.5)) % 8;
    double metric[3], cx[3], cy[3], vx[3], vy[3];
    int64_t kk[3];
    bool any = false;
    for (int i = 0; i < 3; ++i) {
        const int64_t k = ind0 + off[ai + i];
        kk[i] = k;
        cx[i] = F.px[k];
        cy[i] = F.py[k];
        vx[i] = vy[i] = 0.0;
        if (F.sdd[k] < 0) {
            const double ddx = cx[i] - p0x, ddy = cy[i] - p0y;
            const double d = std::sqrt(ddx * ddx + ddy * ddy);
            double dot = n0x * F.vx[k] + n0y * F.vy[k];
            if (dot < 0) {
                vx[i] = -F.vx[k];
                vy[i] = -F.vy[k];
                dot = -dot;
            } else {
                vx[i] = F.vx[k];
                vy[i] = F.vy[k];
            }
            metric[i] = a * d + c * std::acos(dot);  // NaN above 1, as numba's acos
            any = true;
        } else {
            metric[i] = 10000.0;
        }
    }
    if (!any) return false;
    const int b = argmin_np(metric, 3);
    const int64_t k = kk[b];
    if (!(tag[k] + 0.5 < 0)) return false;
    tag[k] = ridge_num;
    p0x = cx[b];
    p0y = cy[b];
    n0x = vx[b];
    n0y = vy[b];
    ind0 = k;
    return true;
}

// _linked_ridge_pts: greedy linking from the strongest unclaimed point.
inline Linked link(const RidgeField& F, double sdd_thresh, double c) {
    const int64_t n = static_cast<int64_t>(F.nx) * F.ny;
    const int64_t ny = F.ny;
    // offsets of (di, dj) = (1,-1), (1,0), (1,1), (0,1), (-1,1), (-1,0),
    // (-1,-1), (0,-1), (1,-1), (1,0) in raveled form
    const int64_t off[10] = {ny - 1, ny, ny + 1, 1, -ny + 1, -ny, -ny - 1, -1, ny - 1, ny};
    const double a = 1 / F.h;
    // The walk visits points in ascending sdd and stops at the first with
    // sdd >= -sdd_thresh, so it only ever sees the points below the
    // threshold, which an ascending order puts first: sort just those (a few
    // percent of the grid), stably, i.e. ties by grid index.
    std::vector<int64_t> order;
    for (int64_t k = 0; k < n; ++k)
        if (F.sdd[k] < -sdd_thresh) order.push_back(k);
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t u, int64_t v) { return F.sdd[u] < F.sdd[v]; });
    std::vector<double> tag(n, -1.0);
    Linked L;
    int64_t start = 0;     // first point of the current ridge in L.pts
    int64_t ridge = 0;     // ridge number counter
    auto count = [&]() { return static_cast<int64_t>(L.pts.size() / 2); };
    for (const int64_t si : order) {
        if (!(tag[si] + 0.5 < 0)) continue;
        if (!(F.sdd[si] < -sdd_thresh)) break;
        tag[si] = static_cast<double>(ridge);
        double p0x = F.px[si], p0y = F.py[si], n0x = F.vx[si], n0y = F.vy[si];
        int64_t ind = si;
        L.pts.push_back(p0x);
        L.pts.push_back(p0y);
        double e0x = p0x, e0y = p0y;
        for (int k = 0; k < 10000; ++k) {
            if (!step(F, tag, p0x, p0y, n0x, n0y, ind, off, static_cast<double>(ridge), a, c))
                break;
            e0x = p0x;
            e0y = p0y;
            L.pts.push_back(p0x);
            L.pts.push_back(p0y);
        }
        // the first walk, reversed, leads into the start point
        for (int64_t lo = start, hi = count() - 1; lo < hi; ++lo, --hi) {
            std::swap(L.pts[2 * lo], L.pts[2 * hi]);
            std::swap(L.pts[2 * lo + 1], L.pts[2 * hi + 1]);
        }
        double p1x = F.px[si], p1y = F.py[si], n1x = -F.vx[si], n1y = -F.vy[si];
        ind = si;
        double e1x = p1x, e1y = p1y;
        for (int k = 0; k < 10000; ++k) {
            if (!step(F, tag, p1x, p1y, n1x, n1y, ind, off, static_cast<double>(ridge), a, c))
                break;
            e1x = p1x;
            e1y = p1y;
            L.pts.push_back(p1x);
            L.pts.push_back(p1y);
        }
        const int64_t cnt = count();
        if (start == cnt - 1) {  // a single point is not a ridge
            L.pts.resize(2 * start);
            continue;
        }
        double t0x = L.pts[2 * (start + 1)] - e0x, t0y = L.pts[2 * (start + 1) + 1] - e0y;
        double s0 = std::sqrt(t0x * t0x + t0y * t0y);
        double t1x = e1x - L.pts[2 * (cnt - 2)], t1y = e1y - L.pts[2 * (cnt - 2) + 1];
        double s1 = std::sqrt(t1x * t1x + t1y * t1y);
        L.end.push_back(cnt);
        L.len.push_back(cnt - start);
        const double r = static_cast<double>(ridge);
        L.ep.insert(L.ep.end(), {e0x, e0y, r, e1x, e1y, -(r + 0.1)});
        L.eptan.insert(L.eptan.end(), {t0x / s0, t0y / s0, t1x / s1, t1y / s1});
        start = cnt;
        ++ridge;
    }
    return L;
}

// Remaining endpoints: rows of (x, y, tag) with their tangents.
struct Endpoints {
    std::vector<double> e;  // 3 per row
    std::vector<double> t;  // 2 per row
    int64_t rows() const { return static_cast<int64_t>(e.size() / 3); }
    void erase2(int64_t i, int64_t j) {  // remove rows i and j
        if (i < j) std::swap(i, j);
        for (int64_t r : {i, j}) {
            e.erase(e.begin() + 3 * r, e.begin() + 3 * r + 3);
            t.erase(t.begin() + 2 * r, t.begin() + 2 * r + 2);
        }
    }
};

inline double clamped_acos(double v) { return std::acos(std::min(1.0, std::max(-1.0, v))); }

inline double sgn(double v) { return std::copysign(1.0, v); }

// _connect_endpoints: the closest aligned endpoint within dist_tol of
// (ex, ey, etag) with tangent (tx, ty).  On success removes that endpoint
// and its partner, returns true and moves (ex, ey, etag, tx, ty) to the
// partner; connect receives the matched tag.
inline bool connect(Endpoints& R, double& ex, double& ey, double& etag, double& tx, double& ty,
                    double ang, double tol, double& connect_tag) {
    const int64_t m = R.rows();
    std::vector<double> dist(m), ux(m), uy(m);
    int64_t n_in = 0;
    for (int64_t k = 0; k < m; ++k) {
        const double ddx = R.e[3 * k] - ex, ddy = R.e[3 * k + 1] - ey;
        dist[k] = std::sqrt(ddx * ddx + ddy * ddy);
        ux[k] = ddx;
        uy[k] = ddy;
        if (dist[k] < tol) {
            ++n_in;
            const double s = std::sqrt(ddx * ddx + ddy * ddy);
            ux[k] = ddx / s;
            uy[k] = ddy / s;
        }
    }
    if (n_in == 0) return false;
    const double es = sgn(etag);
    for (int64_t it = 0; it < n_in; ++it) {
        const int64_t k = argmin_np(dist.data(), static_cast<int>(m));
        const double ns = sgn(R.e[3 * k + 2]);
        const double ntx = R.t[2 * k], nty = R.t[2 * k + 1];
        if (clamped_acos(-es * (ux[k] * tx + uy[k] * ty)) < ang &&
            clamped_acos(ns * (ntx * ux[k] + nty * uy[k])) < ang) {
            const double tagk = R.e[3 * k + 2];
            const int64_t p = k + static_cast<int64_t>(sgn(tagk));
            connect_tag = tagk;
            ex = R.e[3 * p];
            ey = R.e[3 * p + 1];
            etag = R.e[3 * p + 2];
            tx = R.t[2 * p];
            ty = R.t[2 * p + 1];
            R.erase2(k, p);
            return true;
        }
        dist[k] = 10 * tol;
    }
    return false;
}

// Chain from one endpoint of the current ridge; returns the tags, the first
// being -(tag + 0.1) of the starting endpoint.
inline std::vector<double> chain(Endpoints& R, const double* cur, const double* curtan,
                                 double ang, double tol, int64_t nridges) {
    std::vector<double> tags{-(cur[2] + 0.1)};
    double ex = cur[0], ey = cur[1], etag = cur[2], tx = curtan[0], ty = curtan[1];
    for (int64_t k = 0; k < nridges; ++k) {
        double ct;
        if (!connect(R, ex, ey, etag, tx, ty, ang, tol, ct)) break;
        tags.push_back(ct);
    }
    return tags;
}

// Append the curve made of the pieces in tags (negative tag: piece reversed).
inline void assemble(const Linked& L, const std::vector<double>& tags, std::vector<double>& out,
                     std::vector<int64_t>& offsets) {
    for (double tg : tags) {
        const int64_t r = static_cast<int64_t>(std::llround(std::fabs(tg)));
        const int64_t e = L.end[r], l = L.len[r];
        if (tg + 0.01 < 0) {
            for (int64_t q = e - 1; q >= e - l; --q) out.insert(out.end(), {L.pts[2 * q], L.pts[2 * q + 1]});
        } else {
            for (int64_t q = e - l; q < e; ++q) out.insert(out.end(), {L.pts[2 * q], L.pts[2 * q + 1]});
        }
    }
    offsets.push_back(static_cast<int64_t>(out.size() / 2));
}

inline int64_t total_len(const Linked& L, const std::vector<double>& tags) {
    int64_t s = 0;
    for (double tg : tags) s += L.len[static_cast<int64_t>(std::llround(std::fabs(tg)))];
    return s;
}

// ftle_ordered_ridges' joining loop over the linked ridges, scanning every
// remaining endpoint at each query (the restatement's own form).
inline void join_scan(const Linked& L, double tol, double ang, int64_t min_pts,
                      std::vector<double>& out, std::vector<int64_t>& offsets) {
    Endpoints R{L.ep, L.eptan};
    const int64_t nridges = static_cast<int64_t>(L.end.size());
    offsets.assign(1, 0);
    while (R.rows() > 0) {
        const double cur[6] = {R.e[0], R.e[1], R.e[2], R.e[3], R.e[4], R.e[5]};
        const double ctan[4] = {R.t[0], R.t[1], R.t[2], R.t[3]};
        R.erase2(0, 1);
        const int64_t m = R.rows();
        double mind[2] = {INFINITY, INFINITY};
        bool near[2] = {false, false};
        for (int s = 0; s < 2; ++s)
            for (int64_t k = 0; k < m; ++k) {
                const double ddx = R.e[3 * k] - cur[3 * s], ddy = R.e[3 * k + 1] - cur[3 * s + 1];
                const double d = std::sqrt(ddx * ddx + ddy * ddy);
                mind[s] = std::min(mind[s], d);
                near[s] = near[s] || d < tol;
            }
        const int which = (near[0] ? 1 : 0) + (near[1] ? 2 : 0);
        std::vector<double> tags;
        bool check_len = true;
        if (which == 3) {
            const int s0 = mind[1] < mind[0] ? 1 : 0;  // argmin of the two minima
            const std::vector<double> t0 = chain(R, cur + 3 * s0, ctan + 2 * s0, ang, tol, nridges);
            const int s1 = s0 + static_cast<int>(sgn(cur[3 * s0 + 2]));
            const std::vector<double> t1 = chain(R, cur + 3 * s1, ctan + 2 * s1, ang, tol, nridges);
            if (t0.size() > 1 && t1.size() > 1) {
                for (auto it = t0.rbegin(); it != t0.rend(); ++it) tags.push_back(-(*it + 0.1));
                tags.insert(tags.end(), t1.begin() + 1, t1.end());
            } else if (t0.size() > 1) {
                tags = t0;
            } else {
                tags = t1;
            }
            if (t0.size() + t1.size() <= 2) tags.resize(1);
        } else if (which == 2 || which == 1) {
            const int s = which == 2 ? 1 : 0;
            tags = chain(R, cur + 3 * s, ctan + 2 * s, ang, tol, nridges);
            // the start-endpoint case keeps an unjoined curve of any length
            if (which == 1 && tags.size() == 1) check_len = false;
        } else {
            tags = {cur[2]};
        }
        if (check_len && total_len(L, tags) < min_pts) continue;
        if (tags.size() == 1) {
            // one piece, in stored order
            const int64_t r = static_cast<int64_t>(std::llround(std::fabs(tags[0])));
            for (int64_t q = L.end[r] - L.len[r]; q < L.end[r]; ++q)
                out.insert(out.end(), {L.pts[2 * q], L.pts[2 * q + 1]});
            offsets.push_back(static_cast<int64_t>(out.size() / 2));
        } else {
            assemble(L, tags, out, offsets);
        }
    }
}

// The same joining loop with the endpoints in a uniform grid of cell size
// tol.  A query returns the live endpoints within tol (the only ones the
// scan's argmin can pick: all others are at tol or beyond, rejected ones are
// set to 10 tol), and trying them in (distance, row) order is the scan's
// argmin sequence with its first-row tie rule; removals keep the rows'
// relative order, so original indices order them as the scan's rows.  The
// cost is O(endpoints x neighbours) instead of O(endpoints^2).
struct EndpointGrid {
    const std::vector<double>& e;  // 3 per row: x, y, tag
    const std::vector<double>& t;  // 2 per row
    std::vector<char> live;
    double x0 = 0, y0 = 0, cs = 1;
    int64_t ncx = 0, ncy = 0, first = 0;
    std::vector<int64_t> start, ids;  // cells in CSR form, rows ascending per cell

    struct Hit {
        int64_t k;
        double d, dx, dy;
    };

    EndpointGrid(const std::vector<double>& e_, const std::vector<double>& t_) : e(e_), t(t_) {}

    // false when the grid cannot stand in for the scan (non-finite data or a
    // non-positive tolerance).  Cells are at least tol wide (so a point
    // within tol of a query lies in the query's cell or a neighbour, one
    // more for rounding) and doubled until there are at most ~2 per endpoint.
    bool build(double tol) {
        const int64_t m = static_cast<int64_t>(e.size() / 3);
        if (!(tol > 0) || !std::isfinite(tol)) return false;
        double x1 = 0, y1 = 0;
        x0 = y0 = 0;
        for (int64_t k = 0; k < m; ++k) {
            const double x = e[3 * k], y = e[3 * k + 1];
            if (!std::isfinite(x) || !std::isfinite(y)) return false;
            if (k == 0 || x < x0) x0 = x;
            if (k == 0 || y < y0) y0 = y;
            if (k == 0 || x > x1) x1 = x;
            if (k == 0 || y > y1) y1 = y;
        }
        cs = tol;
        double fx = std::floor((x1 - x0) / cs) + 1, fy = std::floor((y1 - y0) / cs) + 1;
        while (!(fx * fy <= 2.0 * m + 16)) {
            cs *= 2;
            fx = std::floor((x1 - x0) / cs) + 1;
            fy = std::floor((y1 - y0) / cs) + 1;
        }
        ncx = static_cast<int64_t>(fx);
        ncy = static_cast<int64_t>(fy);
        std::vector<int64_t> cell(static_cast<size_t>(m));
        start.assign(static_cast<size_t>(ncx * ncy + 1), 0);
        for (int64_t k = 0; k < m; ++k) {
            cell[k] = cell_of(e[3 * k], e[3 * k + 1]);
            ++start[cell[k] + 1];
        }
        for (int64_t c = 0; c < ncx * ncy; ++c) start[c + 1] += start[c];
        ids.assign(static_cast<size_t>(m), 0);
        std::vector<int64_t> fill(start.begin(), start.end() - 1);
        for (int64_t k = 0; k < m; ++k) ids[fill[cell[k]]++] = k;
        live.assign(static_cast<size_t>(m), 1);
        first = 0;
        return true;
    }
    int64_t cell_of(double x, double y) const {
        const int64_t a = std::min<int64_t>(
            ncx - 1, std::max<int64_t>(0, static_cast<int64_t>(std::floor((x - x0) / cs))));
        const int64_t b = std::min<int64_t>(
            ncy - 1, std::max<int64_t>(0, static_cast<int64_t>(std::floor((y - y0) / cs))));
        return a * ncy + b;
    }
    bool empty() {
        const int64_t m = static_cast<int64_t>(live.size());
        while (first < m && !live[first]) ++first;
        return first >= m;
    }
    // live endpoints strictly within tol of (qx, qy), distances as the scan
    // computes them; 5x5 cells so that rounding at a cell edge loses none
    void query(double qx, double qy, double tol, std::vector<Hit>& hits) const {
        hits.clear();
        const double fx = std::floor((qx - x0) / cs), fy = std::floor((qy - y0) / cs);
        if (!std::isfinite(fx) || !std::isfinite(fy)) return;
        const int64_t ix = static_cast<int64_t>(std::max(-4.0, std::min(fx, double(ncx) + 3)));
        const int64_t iy = static_cast<int64_t>(std::max(-4.0, std::min(fy, double(ncy) + 3)));
        for (int64_t a = std::max<int64_t>(0, ix - 2); a <= std::min<int64_t>(ncx - 1, ix + 2); ++a)
            for (int64_t b = std::max<int64_t>(0, iy - 2); b <= std::min<int64_t>(ncy - 1, iy + 2); ++b) {
                const int64_t c = a * ncy + b;
                for (int64_t q = start[c]; q < start[c + 1]; ++q) {
                    const int64_t k = ids[q];
                    if (!live[k]) continue;
                    const double ddx = e[3 * k] - qx, ddy = e[3 * k + 1] - qy;
                    const double d = std::sqrt(ddx * ddx + ddy * ddy);
                    if (d < tol) hits.push_back({k, d, ddx, ddy});
                }
            }
    }
};

inline bool connect_grid(EndpointGrid& R, double& ex, double& ey, double& etag, double& tx,
                         double& ty, double ang, double tol, double& connect_tag,
                         std::vector<EndpointGrid::Hit>& hits) {
    R.query(ex, ey, tol, hits);
    if (hits.empty()) return false;
    std::sort(hits.begin(), hits.end(), [](const EndpointGrid::Hit& u, const EndpointGrid::Hit& v) {
        return u.d < v.d || (u.d == v.d && u.k < v.k);
    });
    const double es = sgn(etag);
    for (const auto& h : hits) {
        const int64_t k = h.k;
        const double s = std::sqrt(h.dx * h.dx + h.dy * h.dy);
        const double ux = h.dx / s, uy = h.dy / s;
        const double ns = sgn(R.e[3 * k + 2]);
        const double ntx = R.t[2 * k], nty = R.t[2 * k + 1];
        if (clamped_acos(-es * (ux * tx + uy * ty)) < ang &&
            clamped_acos(ns * (ntx * ux + nty * uy)) < ang) {
            const double tagk = R.e[3 * k + 2];
            const int64_t p = k + static_cast<int64_t>(sgn(tagk));
            connect_tag = tagk;
            ex = R.e[3 * p];
            ey = R.e[3 * p + 1];
            etag = R.e[3 * p + 2];
            tx = R.t[2 * p];
            ty = R.t[2 * p + 1];
            R.live[k] = 0;
            R.live[p] = 0;
            return true;
        }
    }
    return false;
}

inline std::vector<double> chain_grid(EndpointGrid& R, const double* cur, const double* curtan,
                                      double ang, double tol, int64_t nridges,
                                      std::vector<EndpointGrid::Hit>& hits) {
    std::vector<double> tags{-(cur[2] + 0.1)};
    double ex = cur[0], ey = cur[1], etag = cur[2], tx = curtan[0], ty = curtan[1];
    for (int64_t k = 0; k < nridges; ++k) {
        double ct;
        if (!connect_grid(R, ex, ey, etag, tx, ty, ang, tol, ct, hits)) break;
        tags.push_back(ct);
    }
    return tags;
}

inline void join(const Linked& L, double tol, double ang, int64_t min_pts, std::vector<double>& out,
                 std::vector<int64_t>& offsets) {
    EndpointGrid R(L.ep, L.eptan);
    if (!R.build(tol)) {
        join_scan(L, tol, ang, min_pts, out, offsets);
        return;
    }
    const int64_t nridges = static_cast<int64_t>(L.end.size());
    std::vector<EndpointGrid::Hit> hits;
    offsets.assign(1, 0);
    while (!R.empty()) {
        const int64_t i0 = R.first;  // a pair's start row: pairs leave together
        const double cur[6] = {L.ep[3 * i0],     L.ep[3 * i0 + 1], L.ep[3 * i0 + 2],
                               L.ep[3 * i0 + 3], L.ep[3 * i0 + 4], L.ep[3 * i0 + 5]};
        const double ctan[4] = {L.eptan[2 * i0], L.eptan[2 * i0 + 1], L.eptan[2 * i0 + 2],
                                L.eptan[2 * i0 + 3]};
        R.live[i0] = 0;
        R.live[i0 + 1] = 0;
        double mind[2] = {INFINITY, INFINITY};
        bool near[2] = {false, false};
        for (int s = 0; s < 2; ++s) {
            R.query(cur[3 * s], cur[3 * s + 1], tol, hits);
            near[s] = !hits.empty();
            for (const auto& h : hits) mind[s] = std::min(mind[s], h.d);
        }
        const int which = (near[0] ? 1 : 0) + (near[1] ? 2 : 0);
        std::vector<double> tags;
        bool check_len = true;
        if (which == 3) {
            const int s0 = mind[1] < mind[0] ? 1 : 0;
            const std::vector<double> t0 =
                chain_grid(R, cur + 3 * s0, ctan + 2 * s0, ang, tol, nridges, hits);
            const int s1 = s0 + static_cast<int>(sgn(cur[3 * s0 + 2]));
            const std::vector<double> t1 =
                chain_grid(R, cur + 3 * s1, ctan + 2 * s1, ang, tol, nridges, hits);
            if (t0.size() > 1 && t1.size() > 1) {
                for (auto it = t0.rbegin(); it != t0.rend(); ++it) tags.push_back(-(*it + 0.1));
                tags.insert(tags.end(), t1.begin() + 1, t1.end());
            } else if (t0.size() > 1) {
                tags = t0;
            } else {
                tags = t1;
            }
            if (t0.size() + t1.size() <= 2) tags.resize(1);
        } else if (which == 2 || which == 1) {
            const int s = which == 2 ? 1 : 0;
            tags = chain_grid(R, cur + 3 * s, ctan + 2 * s, ang, tol, nridges, hits);
            if (which == 1 && tags.size() == 1) check_len = false;
        } else {
            tags = {cur[2]};
        }
        if (check_len && total_len(L, tags) < min_pts) continue;
        if (tags.size() == 1) {
            const int64_t r = static_cast<int64_t>(std::llround(std::fabs(tags[0])));
            for (int64_t q = L.end[r] - L.len[r]; q < L.end[r]; ++q)
                out.insert(out.end(), {L.pts[2 * q], L.pts[2 * q + 1]});
            offsets.push_back(static_cast<int64_t>(out.size() / 2));
        } else {
            assemble(L, tags, out, offsets);
        }
    }
}

}  // namespace ordered
}  // namespace nbx
