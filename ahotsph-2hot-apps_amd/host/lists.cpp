// Sink groups and per-group interaction lists (host).
//
// A sink group is a contiguous run of Morton-sorted bodies: the largest cells
// holding <= group_max bodies, with adjacent small siblings merged up to
// group_max so that a wave64 has few idle lanes (nln/mac.c:293-339 evaluates a
// group as the body range [FirstBody, LastBody] under a sink cell).
//
// Each group walks the tree once.  A cell is accepted as a multipole source
// when it is farther than bmax/theta + r_group from the group centre, i.e. the
// opening-angle test holds for every sink of the group (the Bmax MAC of
// nln/cofm.c:568-569 with the sink extent of the dual walk, nln/mac.c:1787-1844).
// A cell that fails the test is opened; a leaf that fails, or a cell with fewer
// bodies than the cost cut, contributes its bodies to the group's body list
// (the terminal-traversal optimisation of nln/mac.c:1834-1843).  Every sink
// therefore interacts with every body exactly once, either directly or inside
// one accepted cell: sum over a group's sources of their body counts == N (the
// invariant nln/mac.c:734 checks).
#include <math.h>

#include <algorithm>
#include <atomic>
#include <functional>
#include <numeric>
#include <stdexcept>
#include <thread>

#include "pgrav_host.h"

namespace pgrav {

static void emit_group(const Tree &t, int64_t first, int64_t count, int gmax, Lists &l) {
    // a leaf at the maximum depth may hold more than gmax coincident bodies
    for (int64_t a = first; a < first + count; a += gmax) {
        l.gfirst.push_back(a);
        l.gcount.push_back((int32_t)std::min<int64_t>(gmax, first + count - a));
    }
}

static void collect(const Tree &t, int32_t nd, const TreeParams &p, Lists &l) {
    const int gmax = p.group_max;
    if (t.count[nd] <= gmax || t.child[nd] < 0) {
        emit_group(t, t.first[nd], t.count[nd], gmax, l);
        return;
    }
    int64_t pf = -1, pc = 0;  // pending merged run of small siblings
    for (int j = 0; j < t.nchild[nd]; j++) {
        const int32_t c = t.child[nd] + j;
        const int32_t cc = t.count[c];
        if (cc <= gmax) {
            if (pf >= 0 && p.merge_groups && pc + cc <= gmax) {
                pc += cc;
                continue;
            }
            if (pf >= 0)
                emit_group(t, pf, pc, gmax, l);
            pf = t.first[c];
            pc = cc;
        } else {
            if (pf >= 0)
                emit_group(t, pf, pc, gmax, l);
            pf = -1;
            pc = 0;
            collect(t, c, p, l);
        }
    }
    if (pf >= 0)
        emit_group(t, pf, pc, gmax, l);
}

static int auto_pcut(int order) {
    // body-count at which one cell interaction (~110 / ~38 VALU for orders 4 / 2)
    // costs no more than its bodies (~15 VALU each)
    return order >= 4 ? 8 : (order >= 2 ? 3 : 2);
}

// The acceptance test and the walk of one sink group (centre g, radius rg).
struct Mac {
    const Tree &t;
    int pcut, kind;
    double inv_theta, box, r_quad, r_mono;
    bool periodic, mixed, arel, boxd;
    int top, quad_ncut, hexa_ncut;
    Mac(const Tree &tree, const TreeParams &p) : t(tree) {
        pcut = p.pcut > 0 ? p.pcut : auto_pcut(p.order);
        kind = p.order >= 4 ? L_HEXA : (p.order >= 2 ? L_QUAD : L_MONO);
        inv_theta = 1.0 / p.theta;
        box = p.box;
        periodic = p.periodic == 1;  // explicit images (2) shift the group instead
        // mixed order: distance (in bmax) beyond which order 2 / order 0 meets the
        // truncation bound theta^(order+1) that the top order meets at bmax/theta
        top = p.order >= 4 ? 4 : (p.order >= 2 ? 2 : 0);
        arel = p.mac == 1;
        boxd = p.group_box == 1;
        if (arel && t.rcrit.size() != 3 * (size_t)t.nnodes())
            throw std::runtime_error("tolerance MAC without critical radii");
        quad_ncut = p.quad_ncut;
        hexa_ncut = p.hexa_ncut;
        mixed = p.mixed_order && top > 0 && !arel;
        r_quad = top == 4 ? pow(p.theta, -5.0 / 3.0) : 1e300;
        r_mono = pow(p.theta, -(double)(top + 1));
    }
    // cell(kind, node) for every accepted cell, bodies(first, count) for every run of
    // bodies listed directly; returns the number of bodies covered (== N when whole).
    // gc: the group's box centre and enclosing radius; gh: the box half-extents, or
    // null to measure from the enclosing sphere
    template <class C, class B>
    int64_t walk(const double *gc, const double *gh, std::vector<int32_t> &stack, C cell,
                 B bodies) const {
        const double gx = gc[0], gy = gc[1], gz = gc[2], rg = gc[3];
        int64_t covered = 0;
        stack.clear();
        stack.push_back(0);
        while (!stack.empty()) {
            const int32_t c = stack.back();
            stack.pop_back();
            double dx = t.cofm[3 * c] - gx, dy = t.cofm[3 * c + 1] - gy, dz = t.cofm[3 * c + 2] - gz;
            if (periodic) {
                dx -= box * nearbyint(dx / box);
                dy -= box * nearbyint(dy / box);
                dz -= box * nearbyint(dz / box);
            }
            // D: a lower bound on the distance from the cell's centre of mass to every
            // sink of the group
            double D;
            if (gh) {
                const double ex = std::max(fabs(dx) - gh[0], 0.0), ey = std::max(fabs(dy) - gh[1], 0.0),
                             ez = std::max(fabs(dz) - gh[2], 0.0);
                D = sqrt(ex * ex + ey * ey + ez * ez);
            } else {
                D = sqrt(dx * dx + dy * dy + dz * dz) - rg;
            }
            const int32_t cnt = t.count[c];
            const double b = t.bmax[c];
            int k = -1;  // the list kind the cell is taken at; -1: not taken
            double reach = 0.0;
            if (arel) {
                // nln/mac.c:1784-1800 for the group's nearest possible sink: order 0 if
                // it clears rcrit + bmax, else order 2, else order 4, each by its radius
                const double *rc = &t.rcrit[3 * (size_t)c];
                if (D > rc[0] + b) {
                    k = L_MONO;
                    reach = rc[0] + b;
                } else if (top >= 2 && quad_ncut > 0 && cnt >= quad_ncut && D > rc[1] + b) {
                    k = L_QUAD;
                    reach = rc[1] + b;
                } else if (top >= 4 && hexa_ncut > 0 && cnt >= hexa_ncut && D > rc[2] + b) {
                    k = L_HEXA;
                    reach = rc[2] + b;
                }
            } else if (cnt >= pcut && D > b * inv_theta) {
                k = kind;
                reach = b * inv_theta;
                if (mixed) {
                    const double s = D / b;
                    k = s > r_mono ? L_MONO : (s > r_quad ? L_QUAD : kind);
                }
            }
            // in a periodic box only cells well inside half a box have an
            // unambiguous nearest image
            if (k >= 0 && periodic && reach + rg + b > 0.5 * box)
                k = -1;
            if (k >= 0) {
                cell(k, c);
                covered += cnt;
            } else if (t.child[c] < 0 || (!arel && cnt < pcut)) {
                // a leaf that is not taken lists its bodies (nln/mac.c:1816-1832)
                bodies(t.first[c], cnt);
                covered += cnt;
            } else {
                for (int j = t.nchild[c] - 1; j >= 0; j--) stack.push_back(t.child[c] + j);
            }
        }
        return covered;
    }
};

double wave_cost(int64_t hexa, int64_t quad, int64_t mono, int64_t body) {
    // VALU instructions per wave of the batched kernel: 110 per order-4 cell,
    // 38 per order-2 cell, 16 per order-0 cell, 15.25 per body, ~40 per 64-entry
    // tile (fetch, offsets, barriers, compensated fold) and ~60 per group
    auto tiles = [](int64_t n) { return (double)((n + 63) / 64); };
    return 110.0 * hexa + 38.0 * quad + 16.0 * mono + 15.25 * body +
           40.0 * (tiles(hexa) + tiles(quad) + tiles(mono) + tiles(body)) + 60.0;
}

// bounding-box centre and enclosing radius of a run of sorted bodies (and, with
// half3, the box's half-extents)
static void group_sphere(const Tree &t, int64_t f, int64_t n, double *out4, double *half3 = nullptr) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t i = f; i < f + n; i++)
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], t.pos[3 * i + a]);
            hi[a] = std::max(hi[a], t.pos[3 * i + a]);
        }
    double r2 = 0;
    for (int a = 0; a < 3; a++) out4[a] = 0.5 * (lo[a] + hi[a]);
    if (half3)
        for (int a = 0; a < 3; a++) half3[a] = 0.5 * (hi[a] - lo[a]);
    for (int64_t i = f; i < f + n; i++) {
        double d2 = 0;
        for (int a = 0; a < 3; a++) d2 += (t.pos[3 * i + a] - out4[a]) * (t.pos[3 * i + a] - out4[a]);
        r2 = std::max(r2, d2);
    }
    out4[3] = sqrt(r2);
}

// bounding-box centre and enclosing radius of a point set (x, y, z triples)
static void sphere_of(const std::vector<double> &p, double *out4, double *half3) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (size_t i = 0; i < p.size(); i += 3)
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], p[i + a]);
            hi[a] = std::max(hi[a], p[i + a]);
        }
    double r2 = 0;
    for (int a = 0; a < 3; a++) out4[a] = 0.5 * (lo[a] + hi[a]);
    for (int a = 0; a < 3; a++) half3[a] = 0.5 * (hi[a] - lo[a]);
    for (size_t i = 0; i < p.size(); i += 3) {
        double d2 = 0;
        for (int a = 0; a < 3; a++) d2 += (p[i + a] - out4[a]) * (p[i + a] - out4[a]);
        r2 = std::max(r2, d2);
    }
    out4[3] = sqrt(r2);
}

static void run_threads(int nth, const std::function<void(int)> &fn) {
    if (nth <= 1) {
        fn(0);
        return;
    }
    std::vector<std::thread> th;
    for (int i = 0; i < nth; i++) th.emplace_back(fn, i);
    for (auto &x : th) x.join();
}

// Cost-driven merging (mode 4): start from sibling-merged cells and merge two
// Morton-adjacent groups when their joint list costs less than the two lists
// (both evaluated with the MAC above); the most profitable disjoint pairs first,
// two rounds.
static void merge_by_cost(const Tree &t, const TreeParams &p, Lists &l) {
    const Mac mac(t, p);
    const int nth = default_threads(p.nthreads);
    auto cost_of = [&](int64_t f, int64_t n, std::vector<int32_t> &stack) {
        double gc[4], gh[3];
        group_sphere(t, f, n, gc, gh);
        int64_t cnt[NLIST] = {0, 0, 0, 0};
        mac.walk(gc, mac.boxd ? gh : nullptr, stack, [&](int k, int32_t) { cnt[k]++; },
                 [&](int64_t, int32_t m) { cnt[L_BODY] += m; });
        return wave_cost(cnt[L_HEXA], cnt[L_QUAD], cnt[L_MONO], cnt[L_BODY]);
    };
    for (int round = 0; round < 2; round++) {
        const int64_t G = (int64_t)l.gfirst.size();
        if (G < 2)
            return;
        std::vector<double> own(G), pair(G, 0.0);
        std::atomic<int64_t> next(0);
        run_threads(nth, [&](int) {
            std::vector<int32_t> stack;
            stack.reserve(256);
            for (;;) {
                const int64_t g0 = next.fetch_add(256);
                if (g0 >= G)
                    break;
                for (int64_t g = g0; g < std::min(G, g0 + 256); g++) {
                    own[g] = cost_of(l.gfirst[g], l.gcount[g], stack);
                    if (g + 1 < G && l.gcount[g] + l.gcount[g + 1] <= p.group_max &&
                        l.gfirst[g] + l.gcount[g] == l.gfirst[g + 1])
                        pair[g] = cost_of(l.gfirst[g], l.gcount[g] + l.gcount[g + 1], stack);
                }
            }
        });
        std::vector<std::pair<double, int64_t>> gain;
        for (int64_t g = 0; g + 1 < G; g++)
            if (pair[g] > 0 && pair[g] < own[g] + own[g + 1])
                gain.push_back({own[g] + own[g + 1] - pair[g], g});
        if (gain.empty())
            return;
        std::sort(gain.begin(), gain.end(), [](const std::pair<double, int64_t> &a,
                                               const std::pair<double, int64_t> &b) {
            return a.first > b.first || (a.first == b.first && a.second < b.second);
        });
        std::vector<char> take(G, 0), used(G, 0);
        for (const auto &x : gain) {
            const int64_t g = x.second;
            if (!used[g] && !used[g + 1]) {
                used[g] = used[g + 1] = 1;
                take[g] = 1;
            }
        }
        std::vector<int64_t> f2;
        std::vector<int32_t> c2;
        for (int64_t g = 0; g < G; g++) {
            if (take[g]) {
                f2.push_back(l.gfirst[g]);
                c2.push_back(l.gcount[g] + l.gcount[g + 1]);
                g++;
            } else {
                f2.push_back(l.gfirst[g]);
                c2.push_back(l.gcount[g]);
            }
        }
        l.gfirst.swap(f2);
        l.gcount.swap(c2);
    }
}

// Balanced recursive coordinate bisection (mode 5): the bodies of a cell are cut
// into m groups of near-equal size; each cut halves the longest extent of the
// members' bounding box at the count that leaves m1 and m - m1 groups on the
// two sides.  Groups come out in the order of the bisection (spatially
// coherent, like a k-d tree walk).
struct RcbOut {
    std::vector<int64_t> slots;
    std::vector<int32_t> count;
};

static void rcb_split(const Tree &t, std::vector<int64_t> &s, size_t a, size_t b, int m, RcbOut &o) {
    if (m <= 1) {
        o.slots.insert(o.slots.end(), s.begin() + a, s.begin() + b);
        o.count.push_back((int32_t)(b - a));
        return;
    }
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (size_t i = a; i < b; i++)
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], t.pos[3 * s[i] + k]);
            hi[k] = std::max(hi[k], t.pos[3 * s[i] + k]);
        }
    int ax = 0;
    for (int k = 1; k < 3; k++)
        if (hi[k] - lo[k] > hi[ax] - lo[ax])
            ax = k;
    const int m1 = m / 2;
    const size_t n1 = a + (size_t)llround((double)(b - a) * m1 / m);
    std::nth_element(s.begin() + a, s.begin() + n1, s.begin() + b, [&](int64_t x, int64_t y) {
        return t.pos[3 * x + ax] < t.pos[3 * y + ax] || (t.pos[3 * x + ax] == t.pos[3 * y + ax] && x < y);
    });
    rcb_split(t, s, a, n1, m1, o);
    rcb_split(t, s, n1, b, m - m1, o);
}

// cells of at most kmax bodies (or leaves), in tree order
static void rcb_cells(const Tree &t, int32_t nd, int64_t kmax, std::vector<int32_t> &out) {
    if (t.count[nd] > kmax && t.child[nd] >= 0) {
        for (int j = 0; j < t.nchild[nd]; j++) rcb_cells(t, t.child[nd] + j, kmax, out);
        return;
    }
    out.push_back(nd);
}

static void rcb_groups(const Tree &t, const TreeParams &p, Lists &l) {
    std::vector<int32_t> cells;
    rcb_cells(t, 0, std::max<int64_t>(p.group_max, p.rcb_cell), cells);
    std::vector<RcbOut> part(cells.size());
    std::atomic<int64_t> next(0);
    run_threads(default_threads(p.nthreads), [&](int) {
        std::vector<int64_t> s;
        for (;;) {
            const int64_t c = next.fetch_add(1);
            if (c >= (int64_t)cells.size())
                break;
            const int32_t nd = cells[c];
            const int64_t n = t.count[nd];
            s.resize(n);
            for (int64_t i = 0; i < n; i++) s[i] = t.first[nd] + i;
            rcb_split(t, s, 0, (size_t)n, (int)((n + p.group_max - 1) / p.group_max), part[c]);
        }
    });
    l.gslot_off.assign(1, 0);
    for (const RcbOut &o : part) {
        for (int32_t c : o.count) {
            l.gfirst.push_back(-1);
            l.gcount.push_back(c);
            l.gslot_off.push_back(l.gslot_off.back() + c);
        }
        l.gslots.insert(l.gslots.end(), o.slots.begin(), o.slots.end());
    }
}

void build_groups(const Tree &t, const TreeParams &p, Lists &l) {
    if (p.group_max < 1 || p.group_max > 64)
        throw std::runtime_error("group_max must be in [1, 64] (one wave64 lane per sink)");
    l = Lists();
    if (p.merge_groups == 2) {
        // fixed runs of group_max consecutive bodies along the Morton curve
        for (int64_t a = 0; a < t.n; a += p.group_max) {
            l.gfirst.push_back(a);
            l.gcount.push_back((int32_t)std::min<int64_t>(p.group_max, t.n - a));
        }
    } else if (p.merge_groups == 3) {
        // maximal cells with <= group_max bodies, then consecutive ones (along the
        // Morton curve, across parents) merged while the merged bounding box
        // stays within twice the larger box edge
        TreeParams q = p;
        q.merge_groups = 0;
        Lists base;
        collect(t, 0, q, base);
        auto box = [&](int64_t f, int64_t c, double *lo, double *hi) {
            for (int a = 0; a < 3; a++) {
                lo[a] = 1e300;
                hi[a] = -1e300;
            }
            for (int64_t i = f; i < f + c; i++)
                for (int a = 0; a < 3; a++) {
                    lo[a] = std::min(lo[a], t.pos[3 * i + a]);
                    hi[a] = std::max(hi[a], t.pos[3 * i + a]);
                }
        };
        auto edge = [](const double *lo, const double *hi) {
            return std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
        };
        int64_t cf = base.gfirst[0], cc = base.gcount[0];
        double clo[3], chi[3];
        box(cf, cc, clo, chi);
        for (size_t k = 1; k < base.gfirst.size(); k++) {
            const int64_t f = base.gfirst[k], c = base.gcount[k];
            double lo[3], hi[3];
            box(f, c, lo, hi);
            double mlo[3], mhi[3];
            for (int a = 0; a < 3; a++) {
                mlo[a] = std::min(lo[a], clo[a]);
                mhi[a] = std::max(hi[a], chi[a]);
            }
            if (cc + c <= p.group_max &&
                edge(mlo, mhi) <= 2.0 * std::max(edge(lo, hi), edge(clo, chi))) {
                cc += c;
                for (int a = 0; a < 3; a++) {
                    clo[a] = mlo[a];
                    chi[a] = mhi[a];
                }
            } else {
                l.gfirst.push_back(cf);
                l.gcount.push_back((int32_t)cc);
                cf = f;
                cc = c;
                for (int a = 0; a < 3; a++) {
                    clo[a] = lo[a];
                    chi[a] = hi[a];
                }
            }
        }
        l.gfirst.push_back(cf);
        l.gcount.push_back((int32_t)cc);
    } else if (p.merge_groups == 5) {
        // cells of <= rcb_cell bodies cut into ceil(n / group_max) balanced groups by
        // coordinate bisection; groups list their bodies (gslots), not a run
        rcb_groups(t, p, l);
    } else if (p.merge_groups == 4) {
        TreeParams q = p;
        q.merge_groups = 1;
        collect(t, 0, q, l);
        merge_by_cost(t, p, l);
    } else {
        collect(t, 0, p, l);
    }
    l.ngroups = (int64_t)l.gfirst.size();
    l.gcenter.assign(4 * l.ngroups, 0.0);
    l.ghalf.assign(3 * l.ngroups, 0.0);
    for (int64_t g = 0; g < l.ngroups; g++) {
        if (l.gfirst[g] >= 0) {
            group_sphere(t, l.gfirst[g], l.gcount[g], &l.gcenter[4 * g], &l.ghalf[3 * g]);
        } else {
            std::vector<double> pts;
            for (int64_t i = l.gslot_off[g]; i < l.gslot_off[g + 1]; i++)
                for (int k = 0; k < 3; k++) pts.push_back(t.pos[3 * l.gslots[i] + k]);
            sphere_of(pts, &l.gcenter[4 * g], &l.ghalf[3 * g]);
        }
    }
}

void build_lists(const Tree &t, const TreeParams &p, Lists &l) {
    const int64_t G = l.ngroups;
    const int nth = default_threads(p.nthreads);
    const Mac mac(t, p);
    struct Buf {
        std::vector<uint32_t> v[NLIST];
        std::vector<uint8_t> im[NLIST];  // explicit images: the entry's image
        std::vector<int64_t> gid;
        std::vector<int64_t> start;  // NLIST per processed group
        int64_t pairs[NLIST] = {0, 0, 0, 0};
    };
    const bool images = p.periodic == 2;
    if (images && (p.nimage < 0 || p.nimage > MAX_NIMAGE))
        throw std::runtime_error("nimage must be 0, 1 or 2 for explicit periodic images");
    const int h = images ? p.nimage : 0, f = 2 * h + 1, nim = f * f * f;
    std::vector<Buf> bufs(nth);
    std::atomic<int64_t> next(0);
    std::atomic<bool> miscount(false);
    const int64_t chunk = 64;
    auto worker = [&](int tid) {
        Buf &B = bufs[tid];
        std::vector<int32_t> stack;
        stack.reserve(256);
        for (;;) {
            const int64_t g0 = next.fetch_add(chunk);
            if (g0 >= G)
                break;
            for (int64_t g = g0; g < std::min(G, g0 + chunk); g++) {
                B.gid.push_back(g);
                for (int k = 0; k < NLIST; k++) B.start.push_back((int64_t)B.v[k].size());
                // explicit images (nln/mac.c:91-111): the group walks the tree once per
                // image offset o, as if the sources sat at x + o; entries carry the image
                int64_t covered = 0;
                for (int im = 0; im < nim; im++) {
                    const double o[3] = {(im / (f * f) - h) * p.box, ((im / f) % f - h) * p.box,
                                         (im % f - h) * p.box};
                    const double gc[4] = {l.gcenter[4 * g] - o[0], l.gcenter[4 * g + 1] - o[1],
                                          l.gcenter[4 * g + 2] - o[2], l.gcenter[4 * g + 3]};
                    covered += mac.walk(
                        gc, mac.boxd ? &l.ghalf[3 * g] : nullptr, stack,
                        [&](int k, int32_t c) {
                            B.v[k].push_back((uint32_t)c);
                            if (images)
                                B.im[k].push_back((uint8_t)im);
                        },
                        [&](int64_t f0, int32_t n) {
                            for (int64_t i = f0; i < f0 + n; i++) B.v[L_BODY].push_back((uint32_t)i);
                            if (images)
                                B.im[L_BODY].insert(B.im[L_BODY].end(), (size_t)n, (uint8_t)im);
                        });
                }
                if (covered != t.n * nim)  // nln/mac.c:734: Nimage * N interactions per sink
                    miscount.store(true);
                const int64_t *st = &B.start[B.start.size() - NLIST];
                for (int k = 0; k < NLIST; k++)
                    B.pairs[k] += (int64_t)l.gcount[g] * ((int64_t)B.v[k].size() - st[k]);
            }
        }
    };
    if (nth <= 1) {
        worker(0);
    } else {
        std::vector<std::thread> th;
        for (int i = 0; i < nth; i++) th.emplace_back(worker, i);
        for (auto &x : th) x.join();
    }
    if (miscount.load())
        throw std::runtime_error("a group's sources do not cover N bodies per image (nln/mac.c:734)");
    // assemble CSR in group order
    std::vector<int32_t> owner(G);
    std::vector<int64_t> slot(G);
    for (int tid = 0; tid < nth; tid++)
        for (size_t j = 0; j < bufs[tid].gid.size(); j++) {
            owner[bufs[tid].gid[j]] = tid;
            slot[bufs[tid].gid[j]] = (int64_t)j;
        }
    for (int k = 0; k < NLIST; k++) {
        l.off[k].assign(G + 1, 0);
        l.pairs[k] = 0;
    }
    auto seg = [&](int64_t g, int k, int64_t &a, int64_t &b) {
        const Buf &B = bufs[owner[g]];
        const int64_t j = slot[g];
        a = B.start[j * NLIST + k];
        b = (j + 1 < (int64_t)B.gid.size()) ? B.start[(j + 1) * NLIST + k] : (int64_t)B.v[k].size();
    };
    for (int64_t g = 0; g < G; g++)
        for (int k = 0; k < NLIST; k++) {
            int64_t a, b;
            seg(g, k, a, b);
            l.off[k][g + 1] = l.off[k][g] + (b - a);
        }
    for (int k = 0; k < NLIST; k++) {
        l.idx[k].resize(l.off[k][G]);
        l.img[k].resize(images ? l.off[k][G] : 0);
        std::vector<std::thread> th;
        const int nt = std::min(nth, 16);
        for (int t0 = 0; t0 < nt; t0++)
            th.emplace_back([&, t0] {
                for (int64_t g = G * t0 / nt; g < G * (t0 + 1) / nt; g++) {
                    int64_t a, b;
                    seg(g, k, a, b);
                    std::copy(bufs[owner[g]].v[k].begin() + a, bufs[owner[g]].v[k].begin() + b,
                              l.idx[k].begin() + l.off[k][g]);
                    if (images)
                        std::copy(bufs[owner[g]].im[k].begin() + a, bufs[owner[g]].im[k].begin() + b,
                                  l.img[k].begin() + l.off[k][g]);
                }
            });
        for (auto &x : th) x.join();
        for (int tid = 0; tid < nth; tid++) l.pairs[k] += bufs[tid].pairs[k];
    }
    bufs.clear();
    // dispatch order from the per-wave VALU estimate
    std::vector<double> cost(G);
    for (int64_t g = 0; g < G; g++)
        cost[g] = wave_cost(l.off[L_HEXA][g + 1] - l.off[L_HEXA][g], l.off[L_QUAD][g + 1] - l.off[L_QUAD][g],
                            l.off[L_MONO][g + 1] - l.off[L_MONO][g], l.off[L_BODY][g + 1] - l.off[L_BODY][g]);
    make_dispatch(cost, p.dispatch, l);
}

void make_dispatch(const std::vector<double> &cost, int mode, Lists &l) {
    const int64_t G = (int64_t)cost.size();
    const int X = HOST_NXCD;
    std::vector<std::vector<int32_t>> slice(X);
    if (mode == 0) {
        // global longest-first, dealt round-robin to the slices
        std::vector<int32_t> all(G);
        std::iota(all.begin(), all.end(), 0);
        std::stable_sort(all.begin(), all.end(), [&](int32_t a, int32_t b) { return cost[a] > cost[b]; });
        for (int64_t i = 0; i < G; i++) slice[i % X].push_back(all[i]);
    } else {
        // contiguous Morton ranges of equal total cost, longest-first inside
        double total = 0;
        for (double c : cost) total += c;
        double acc = 0;
        int s = 0;
        for (int64_t g = 0; g < G; g++) {
            if (acc >= total * (s + 1) / X && s < X - 1)
                s++;
            slice[s].push_back((int32_t)g);
            acc += cost[g];
        }
        if (mode == 1) {
            for (auto &v : slice)
                std::stable_sort(v.begin(), v.end(), [&](int32_t a, int32_t b) { return cost[a] > cost[b]; });
        } else if (mode == 3) {
            // the costliest 2 % of each slice first, the rest in Morton order
            for (auto &v : slice) {
                std::vector<int32_t> byc(v);
                std::stable_sort(byc.begin(), byc.end(), [&](int32_t a, int32_t b) { return cost[a] > cost[b]; });
                const size_t k = byc.size() / 50;
                std::vector<char> head(G, 0);
                std::vector<int32_t> out(byc.begin(), byc.begin() + k);
                for (size_t i = 0; i < k; i++) head[byc[i]] = 1;
                for (int32_t g : v)
                    if (!head[g])
                        out.push_back(g);
                v.swap(out);
            }
        }
        // mode 2: Morton order inside each slice
    }
    l.lpt.clear();
    l.xoff.assign(X + 1, 0);
    for (int x = 0; x < X; x++) {
        l.xoff[x] = (int32_t)l.lpt.size();
        l.lpt.insert(l.lpt.end(), slice[x].begin(), slice[x].end());
    }
    l.xoff[X] = (int32_t)l.lpt.size();
    // one workgroup per group: block b lands on XCD b % X, so deal the slices
    // round-robin (a slice that runs out is skipped)
    l.order_np.clear();
    std::vector<size_t> pos(X, 0);
    for (int64_t b = 0; (int64_t)l.order_np.size() < G; b++) {
        const int x = (int)(b % X);
        if (pos[x] < slice[x].size())
            l.order_np.push_back(slice[x][pos[x]++]);
        else
            for (int y = 0; y < X; y++)
                if (pos[y] < slice[y].size()) {
                    l.order_np.push_back(slice[y][pos[y]++]);
                    break;
                }
    }
}

}  // namespace pgrav
