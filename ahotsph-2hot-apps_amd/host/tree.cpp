// Octree build and FP64 moments on the host.
//
// Bodies are sorted along a 63-bit Morton key (21 bits per axis, the bit
// interleave of nln/physics_generic.c:146-224 GetKey) with a threaded LSD radix
// sort; cells are key-prefix ranges of the sorted array, so every cell's bodies
// are contiguous (what nln/mac.c:293-339 FirstBody/LastBody rely on).  Moments
// are accumulated in FP64 about each cell's centre of mass: directly from the
// bodies for a leaf (nln/cofm.c:650-768 mpole_add_mono) and by translating the
// daughters' moments for an interior cell (nln/cofm.c:770-964 mpole_add), with
// bmax from the daughters and the cell-geometry bound of nln/cofm.c:214-222.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <functional>
#include <cmath>
#include <stdexcept>
#include <string>
#include <thread>

#include "pgrav_host.h"
#include "pgrav_record.h"

namespace pgrav {

int default_threads(int requested) {
    if (requested > 0)
        return requested;
    unsigned h = std::thread::hardware_concurrency();
    return h ? (int)std::min(h, 64u) : 1;
}

// static split of [0, n) over nthreads
void parallel_for(int nthreads, int64_t n, const std::function<void(int64_t, int64_t)> &fn) {
    if (nthreads <= 1 || n < 4096) {
        fn(0, n);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) {
        int64_t a = n * t / nthreads, b = n * (t + 1) / nthreads;
        th.emplace_back([&fn, a, b] { fn(a, b); });
    }
    for (auto &x : th) x.join();
}

// ---------------------------------------------------------------- Morton keys
static inline uint64_t spread3(uint64_t v) {  // 21 bits -> every third bit
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}

static inline uint64_t compact3(uint64_t v) {  // inverse of spread3
    v &= 0x1249249249249249ULL;
    v = (v ^ (v >> 2)) & 0x10c30c30c30c30c3ULL;
    v = (v ^ (v >> 4)) & 0x100f00f00f00f00fULL;
    v = (v ^ (v >> 8)) & 0x1f0000ff0000ffULL;
    v = (v ^ (v >> 16)) & 0x1f00000000ffffULL;
    v = (v ^ (v >> 32)) & 0x1fffff;
    return v;
}

static void radix_sort(std::vector<uint64_t> &key, std::vector<int64_t> &idx, int nthreads) {
    const int64_t n = (int64_t)key.size();
    std::vector<uint64_t> k2(n);
    std::vector<int64_t> i2(n);
    nthreads = std::max(1, std::min(nthreads, (int)std::max<int64_t>(1, n / 65536)));
    std::vector<int64_t> hist((size_t)nthreads * 256);
    for (int pass = 0; pass < 8; pass++) {
        const int sh = pass * 8;
        std::fill(hist.begin(), hist.end(), 0);
        parallel_for(nthreads, nthreads, [&](int64_t t0, int64_t t1) {
            for (int64_t t = t0; t < t1; t++) {
                int64_t a = n * t / nthreads, b = n * (t + 1) / nthreads;
                int64_t *h = &hist[(size_t)t * 256];
                for (int64_t i = a; i < b; i++) h[(key[i] >> sh) & 255]++;
            }
        });
        // exclusive prefix over (digit, thread)
        int64_t sum = 0;
        for (int d = 0; d < 256; d++)
            for (int t = 0; t < nthreads; t++) {
                int64_t c = hist[(size_t)t * 256 + d];
                hist[(size_t)t * 256 + d] = sum;
                sum += c;
            }
        parallel_for(nthreads, nthreads, [&](int64_t t0, int64_t t1) {
            for (int64_t t = t0; t < t1; t++) {
                int64_t a = n * t / nthreads, b = n * (t + 1) / nthreads;
                int64_t *h = &hist[(size_t)t * 256];
                for (int64_t i = a; i < b; i++) {
                    int64_t o = h[(key[i] >> sh) & 255]++;
                    k2[o] = key[i];
                    i2[o] = idx[i];
                }
            }
        });
        key.swap(k2);
        idx.swap(i2);
    }
}

// ---------------------------------------------------------------- moment shift
// Raw-moment exponent table (nln/cofm.c:660-720 order).
struct MomentTable {
    int e[35][3];
    struct Term {
        int beta;
        double coef;
        int dx, dy, dz;
    };
    std::vector<Term> terms[35];
    MomentTable() {
        int k = 0;
        for (int order = 0; order <= 4; order++)
            for (int a = order; a >= 0; a--)
                for (int b = order - a; b >= 0; b--) {
                    e[k][0] = a;
                    e[k][1] = b;
                    e[k][2] = order - a - b;
                    k++;
                }
        auto binom = [](int n, int r) {
            double c = 1;
            for (int i = 0; i < r; i++) c = c * (n - i) / (i + 1);
            return c;
        };
        for (int al = 0; al < 35; al++)
            for (int be = 0; be < 35; be++) {
                if (e[be][0] > e[al][0] || e[be][1] > e[al][1] || e[be][2] > e[al][2])
                    continue;
                double c = binom(e[al][0], e[be][0]) * binom(e[al][1], e[be][1]) *
                           binom(e[al][2], e[be][2]);
                terms[al].push_back({be, c, e[al][0] - e[be][0], e[al][1] - e[be][1],
                                     e[al][2] - e[be][2]});
            }
    }
};
static const MomentTable &mtab() {
    static MomentTable t;
    return t;
}

int raw_index(int a, int b, int c) {
    const MomentTable &T = mtab();
    for (int k = 0; k < 35; k++)
        if (T.e[k][0] == a && T.e[k][1] == b && T.e[k][2] == c)
            return k;
    throw std::runtime_error("raw_index: no such moment");
}

// dst += moments `src` (about c_src) translated to c_dst; d = c_src - c_dst
static void shift_add(const double *src, const double d[3], double *dst) {
    const MomentTable &T = mtab();
    double pw[3][5];
    for (int a = 0; a < 3; a++) {
        pw[a][0] = 1;
        for (int k = 1; k < 5; k++) pw[a][k] = pw[a][k - 1] * d[a];
    }
    for (int al = 0; al < 35; al++) {
        double s = 0;
        for (const auto &t : T.terms[al]) s += t.coef * pw[0][t.dx] * pw[1][t.dy] * pw[2][t.dz] * src[t.beta];
        dst[al] += s;
    }
}

static void body_moments(const double *p, double m, const double c[3], double *raw) {
    const MomentTable &T = mtab();
    const double x = p[0] - c[0], y = p[1] - c[1], z = p[2] - c[2];
    double px[5], py[5], pz[5];
    px[0] = py[0] = pz[0] = 1;
    for (int k = 1; k < 5; k++) {
        px[k] = px[k - 1] * x;
        py[k] = py[k - 1] * y;
        pz[k] = pz[k - 1] * z;
    }
    for (int al = 0; al < 35; al++) raw[al] += m * px[T.e[al][0]] * py[T.e[al][1]] * pz[T.e[al][2]];
}

// ---------------------------------------------------------------- tree
void build_tree(const double *pos, const double *mass, int64_t n, const TreeParams &p, Tree &t) {
    const int nth = default_threads(p.nthreads);
    t = Tree();
    t.n = n;
    if (n <= 0)
        throw std::runtime_error("build_tree: no bodies");
    // every mass positive and finite, every coordinate finite (the reference
    // aborts on a bad particle mass, nln/mac.c:1838-1843)
    for (int64_t i = 0; i < n; i++) {
        if (!(mass[i] > 0) || !std::isfinite(mass[i]))
            throw std::runtime_error("body " + std::to_string(i) + ": mass must be positive and finite");
        for (int a = 0; a < 3; a++)
            if (!std::isfinite(pos[3 * i + a]))
                throw std::runtime_error("body " + std::to_string(i) + ": position is not finite");
    }
    // bounding cube
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    if (p.periodic) {
        for (int a = 0; a < 3; a++) {
            lo[a] = 0;
            hi[a] = p.box;
        }
        t.size = p.box;
    } else {
        for (int64_t i = 0; i < n; i++)
            for (int a = 0; a < 3; a++) {
                lo[a] = std::min(lo[a], pos[3 * i + a]);
                hi[a] = std::max(hi[a], pos[3 * i + a]);
            }
        double s = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
        t.size = s > 0 ? s * (1 + 1e-9) : 1.0;
    }
    for (int a = 0; a < 3; a++) t.x0[a] = lo[a];

    std::vector<uint64_t> key(n);
    t.perm.resize(n);
    const double scale = (double)(1 << 21) / t.size;
    parallel_for(nth, n, [&](int64_t a, int64_t b) {
        for (int64_t i = a; i < b; i++) {
            uint64_t c[3];
            for (int k = 0; k < 3; k++) {
                double x = pos[3 * i + k];
                if (p.periodic)
                    x -= p.box * floor(x / p.box);
                double f = (x - t.x0[k]) * scale;
                int64_t q = (int64_t)f;
                q = std::max<int64_t>(0, std::min<int64_t>((1 << 21) - 1, q));
                c[k] = (uint64_t)q;
            }
            key[i] = spread3(c[0]) << 2 | spread3(c[1]) << 1 | spread3(c[2]);
            t.perm[i] = i;
        }
    });
    radix_sort(key, t.perm, nth);
    t.pos.resize(3 * n);
    t.mass.resize(n);
    parallel_for(nth, n, [&](int64_t a, int64_t b) {
        for (int64_t i = a; i < b; i++) {
            int64_t j = t.perm[i];
            for (int k = 0; k < 3; k++) {
                double x = pos[3 * j + k];
                if (p.periodic)
                    x -= p.box * floor(x / p.box);
                t.pos[3 * i + k] = x;
            }
            t.mass[i] = mass[j];
        }
    });

    // nodes: children are appended contiguously when their parent is split
    struct Pending {
        int32_t node;
    };
    t.first.push_back(0);
    t.count.push_back((int32_t)n);
    t.child.push_back(-1);
    t.nchild.push_back(0);
    t.level.push_back(0);
    t.parent.push_back(-1);
    std::vector<Pending> stack{{0}};
    const int leaf_max = std::max(1, p.leaf_max);
    while (!stack.empty()) {
        int32_t nd = stack.back().node;
        stack.pop_back();
        const int L = t.level[nd];
        if (t.count[nd] <= leaf_max || L >= 21)
            continue;
        const int64_t a = t.first[nd], b = a + t.count[nd];
        const int sh = 3 * (21 - (L + 1));
        int64_t bound[9];
        bound[0] = a;
        for (int d = 1; d < 8; d++) {
            // first body with digit >= d
            int64_t lo2 = bound[d - 1], hi2 = b;
            while (lo2 < hi2) {
                int64_t mid = (lo2 + hi2) / 2;
                if ((int)((key[mid] >> sh) & 7) < d)
                    lo2 = mid + 1;
                else
                    hi2 = mid;
            }
            bound[d] = lo2;
        }
        bound[8] = b;
        int32_t c0 = (int32_t)t.count.size();
        int nc = 0;
        for (int d = 0; d < 8; d++) {
            if (bound[d + 1] == bound[d])
                continue;
            t.first.push_back(bound[d]);
            t.count.push_back((int32_t)(bound[d + 1] - bound[d]));
            t.child.push_back(-1);
            t.nchild.push_back(0);
            t.level.push_back((uint8_t)(L + 1));
            t.parent.push_back(nd);
            nc++;
        }
        t.child[nd] = c0;
        t.nchild[nd] = (uint8_t)nc;
        for (int c = nc - 1; c >= 0; c--) stack.push_back({c0 + c});
    }
    compute_moments(t, nth, &key);
    if (p.mac == 1)
        compute_rcrit(t, p, nth);
}

// Moments, centres of mass and bmax of every node, deepest level first (every
// child index exceeds its parent's).  With `key`, bmax is also bounded by the
// cell geometry; without it (bodies moved after the build) only the daughters'
// bound is used, which stays valid wherever the bodies went.
void compute_moments(Tree &t, int nth, const std::vector<uint64_t> *keyp) {
    const int64_t nn = t.nnodes();
    t.cofm.assign(3 * nn, 0.0);
    t.bmax.assign(nn, 0.0);
    t.raw.assign(35 * nn, 0.0);
    std::vector<std::vector<int32_t>> bylevel(23);
    for (int64_t i = 0; i < nn; i++) bylevel[t.level[i]].push_back((int32_t)i);
    for (int L = 22; L >= 0; L--) {
        const auto &v = bylevel[L];
        parallel_for(nth, (int64_t)v.size(), [&](int64_t a, int64_t b) {
            for (int64_t k = a; k < b; k++) {
                const int32_t nd = v[k];
                double *c = &t.cofm[3 * nd];
                double *raw = &t.raw[35 * (size_t)nd];
                double M = 0;
                if (t.child[nd] < 0) {
                    const int64_t f = t.first[nd];
                    for (int64_t i = f; i < f + t.count[nd]; i++) {
                        M += t.mass[i];
                        for (int q = 0; q < 3; q++) c[q] += t.mass[i] * t.pos[3 * i + q];
                    }
                    for (int q = 0; q < 3; q++) c[q] /= M;
                    double bm = 0;
                    for (int64_t i = f; i < f + t.count[nd]; i++) {
                        body_moments(&t.pos[3 * i], t.mass[i], c, raw);
                        double dx = t.pos[3 * i] - c[0], dy = t.pos[3 * i + 1] - c[1],
                               dz = t.pos[3 * i + 2] - c[2];
                        bm = std::max(bm, sqrt(dx * dx + dy * dy + dz * dz));
                    }
                    t.bmax[nd] = bm;
                } else {
                    const int32_t c0 = t.child[nd];
                    for (int j = 0; j < t.nchild[nd]; j++) {
                        const double m = t.raw[35 * (size_t)(c0 + j)];
                        M += m;
                        for (int q = 0; q < 3; q++) c[q] += m * t.cofm[3 * (c0 + j) + q];
                    }
                    for (int q = 0; q < 3; q++) c[q] /= M;
                    double bm = 0;
                    for (int j = 0; j < t.nchild[nd]; j++) {
                        const int32_t ch = c0 + j;
                        double d[3] = {t.cofm[3 * ch] - c[0], t.cofm[3 * ch + 1] - c[1],
                                       t.cofm[3 * ch + 2] - c[2]};
                        shift_add(&t.raw[35 * (size_t)ch], d, raw);
                        bm = std::max(bm, sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) + t.bmax[ch]);
                    }
                    t.bmax[nd] = bm;
                    if (!keyp)
                        continue;
                    const std::vector<uint64_t> &key = *keyp;
                    // geometric bound (nln/cofm.c:214-222): half cell + |cofm - cell centre|
                    const double cs = t.size / (double)(1u << t.level[nd]);
                    // cell corner from the key prefix of its first body
                    const int L = t.level[nd];
                    const uint64_t kk = L ? key[t.first[nd]] >> (3 * (21 - L)) : 0;
                    const uint64_t ic[3] = {compact3(kk >> 2), compact3(kk >> 1), compact3(kk)};
                    double g2 = 0;
                    for (int q = 0; q < 3; q++) {
                        double centre = t.x0[q] + ((double)ic[q] + 0.5) * cs;
                        double dq = 0.5 * cs + fabs(c[q] - centre);
                        g2 += dq * dq;
                    }
                    t.bmax[nd] = std::min(bm, sqrt(g2));
                }
                if (t.count[nd] == 1)
                    t.bmax[nd] = 0.0;
            }
        });
    }
}

// ---------------------------------------------------------------- records
void prepare_cells(const Tree &t, int order, std::vector<float> &rec32, std::vector<double> *rec64,
                   std::vector<double> &cpos4) {
    const int64_t nn = t.nnodes();
    rec32.assign((size_t)nn * PREC4, 0.0f);
    if (rec64)
        rec64->assign((size_t)nn * PREC4, 0.0);
    cpos4.assign((size_t)nn * 4, 0.0);
    parallel_for(default_threads(0), nn, [&](int64_t a, int64_t b) {
        double q[29], out[PREC4];
        for (int64_t i = a; i < b; i++) {
            const double *raw = &t.raw[35 * (size_t)i];
            // unnormalised record (R = 1): the prepared values do not depend on R
            raw_to_record<double>(raw, &t.cofm[3 * i], 1.0, q);
            q[5] = q[6] = q[7] = 0.0;  // the dipole about the centre of mass is zero
            prep_record<double>(q, order, 0, out);
            out[PREC_TRACE] = raw_trace<double>(raw);
            for (int k = 0; k < PREC4; k++) rec32[(size_t)i * PREC4 + k] = (float)out[k];
            if (rec64)
                for (int k = 0; k < PREC4; k++) (*rec64)[(size_t)i * PREC4 + k] = out[k];
            cpos4[4 * i + 0] = t.cofm[3 * i + 0];
            cpos4[4 * i + 1] = t.cofm[3 * i + 1];
            cpos4[4 * i + 2] = t.cofm[3 * i + 2];
            cpos4[4 * i + 3] = raw[0];
        }
    });
}

}  // namespace pgrav
