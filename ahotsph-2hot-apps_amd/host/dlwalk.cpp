// The reference's isolated-run list builder for the drop-in path: the "Don't
// Laugh" dual walk with DLRcritMAC (nln/mac.c:1773-1844), the sink inheritance
// of InheritSinkNlogN (nln/mac.c:477-872) and the M x N batching of mxn_hexa /
// mxn_quad2 with the pQ1 queue (nln/mac.c:877-1272, :394-475).  It produces what
// the reference hands nln/gravcuda.h: SoA-NSSE blocks of absolute-position
// records (Hvec / Qvec / Mvec, nln/mac.c:117-153, NSSE = 8) and the sequence of
// grav_mn_CUDA("pH" / "pQ") and grav_qnss_CUDA("pQ1_qnss") calls, plus the
// per-body remainders the reference evaluates on the CPU at the leaf
// (nln/mac.c:506-700).
//
// What is the reference's and what is this file's:
//  - dl_mac restates DLRcritMAC per (sink, source vector) call: the acceptance
//    order M -> Q -> H on the distance from the sink's centre, dr2 >
//    (rcrit + bmax)^2 with the source's own bmax; otherwise SPLIT_SINK when
//    dlfac * sink bmax exceeds the source's smallest critical radius, else
//    SPLIT_SRC; body sources wait (SPLIT_SINK) until the sink is a body; and the
//    terminal expansion of a SPLIT_SRC cell into its bodies.  FP32 arithmetic on
//    FP32 cell fields, as the reference's cell structs hold them.
//  - The walk that calls it (WalkNT) lives in the reference's libtree, which is
//    not part of the reference checkout.  This file's walk: depth first over the
//    sink tree in Morton order, one MAC call per sink with every pending source,
//    the SPLIT_SINK sources passed to each daughter, and -- in the gated mode
//    only, where SPLIT_SRC can remain -- the split source's daughters presented
//    to the same sink in a further call.  The root sink takes the root cell's
//    centre and bmax (the reference's from == NULL branch clears them,
//    nln/mac.c:848, leaving the root's decisions to the absent walk).
//  - gated = 0 keeps the reference's terminal expansion of every SPLIT_SRC cell
//    (no leaf_max_n gate, nln/mac.c:1818); gated = 1 expands only cells of at
//    most leaf_max_n bodies (the gate of DLRcritMACsb, nln/mac.c:1740).
//  - The pQ1 queue's tail is flushed at the end of the walk (the reference never
//    flushes it, nln/mac.c:957).
//
// Everything here is host code; the GPU evaluates the emitted calls through the
// drop-in ABI.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "pgrav.h"
#include "pgrav_host.h"
#include "pgrav_record.h"

namespace pgrav {
void fail(const char *fmt, ...);
void clear_error();

namespace {

const int NSSE = 8;
const int HSZ = 29, QSZ = 13, MSZ = 4;  // floats per record (nln/pgrav.cu HSZ / QSZ / MSZ)
enum { MAC_ACCEPT = 0, MAC_SPLIT_SINK = 1, MAC_SPLIT_SRC = 2 };
enum { APP_M = 0, APP_Q = 1, APP_H = 2 };

// A walk entity: a tree node (>= 0) or a body, -(sorted slot + 1).
inline int32_t body_ent(int64_t slot) { return (int32_t)(-(slot + 1)); }
inline int64_t ent_body(int32_t e) { return -(int64_t)e - 1; }

struct Src {
    int32_t ent;
    int32_t img;  // offset index (nln/mac.c:91-111 numbering)
};

// A cell as the reference's cell / quadcell / hexacell structs hold it (FP32).
struct Cell {
    float mass, pos[3], rcrit, bmax, rcrit_q, rcrit_h;
    int64_t daughters;  // bodies below
};

// The reference's Sink (physics.h) fields this walk uses.
struct Sink {
    float pos[3], bmax;
    int64_t daughters;
    bool isbody, processed;
    int64_t first;  // first body slot (FirstBody)
    int64_t mcnt, qcnt, hcnt, qcnt_done, hcnt_done;
    int64_t interactions;
};

// A stack of SoA-NSSE records: record k's field f is at
// [(k / NSSE) * sz * NSSE + f * NSSE + k % NSSE].
struct SoA {
    int sz;
    std::vector<float> v;
    explicit SoA(int s) : sz(s) {}
    float *rec_field(int64_t k, int f) {
        const size_t need = (size_t)(k / NSSE + 1) * sz * NSSE;
        if (v.size() < need)
            v.resize(std::max(need, v.size() * 2), 0.0f);
        return &v[(size_t)(k / NSSE) * sz * NSSE + (size_t)f * NSSE + k % NSSE];
    }
    void put(int64_t k, const float *rec) {
        for (int f = 0; f < sz; f++) *rec_field(k, f) = rec[f];
    }
    // zero records [k0, k1) (the leaf's padding, nln/mac.c:512-548)
    void zero(int64_t k0, int64_t k1) {
        for (int64_t k = k0; k < k1; k++)
            for (int f = 0; f < sz; f++) *rec_field(k, f) = 0.0f;
    }
    const float *block(int64_t b) {
        rec_field(b * NSSE + NSSE - 1, sz - 1);
        return &v[(size_t)b * sz * NSSE];
    }
};

struct DLWalk {
    TreeParams tp;
    pgrav_dl_params dp;
    Tree tree;
    int p2cut = 0, p4cut = 0;
    int nimage_side = 0, nimage = 1;
    std::vector<float> offset;   // 3 per image
    std::vector<Cell> cell;      // per node
    std::vector<float> rec;      // 29 per node: the reference hexacell layout (float)
    std::vector<float> body;     // 4 per sorted slot: mass, x, y, z (float)
    std::vector<int64_t> child_off;
    std::vector<int32_t> children;
    SoA H{HSZ}, Q{QSZ}, M{MSZ};
    // outputs
    std::vector<int64_t> calls;   // 6 per M x N call: routine (0 pH, 1 pQ), first, m, src_off, source_n, sz
    std::vector<float> src;       // every emitted source block
    std::vector<int64_t> qflush;  // 6 per pQ1 flush: sink_base, m, seg_off, src_off, source_len, 0
    std::vector<uint32_t> qseg;   // 2 per queued sink: base, length
    std::vector<int64_t> tails;   // 6 per body slot: h_off, h_n, q_off, q_n, m_off, m_n
    std::vector<int64_t> inter;   // interactions per body slot
    // the pQ1 queue (nln/mac.c:172-184)
    struct {
        bool inuse = false;
        int64_t sink_base = 0, m = 0;
        std::vector<uint32_t> seg;  // 2 per sink
        std::vector<float> source;  // SoA blocks
    } qq;
    // MAC call log
    std::vector<int64_t> log_calls;  // 4 per call: sink ent, src_off, n, app_off
    std::vector<int32_t> log_src;    // 3 per source: ent, img, result
    std::vector<int32_t> log_app;    // 3 per append: kind, ent, img
    int64_t stats[16] = {0};

    // ------------------------------------------------------------ MAC
    void position(int32_t ent, int img, float r[3]) const {
        const float *p = ent >= 0 ? cell[ent].pos : &body[4 * ent_body(ent) + 1];
        const float *o = &offset[3 * img];
        r[0] = p[0] + o[0];
        r[1] = p[1] + o[1];
        r[2] = p[2] + o[2];
    }
    void append(Sink &s, int kind, int32_t ent, int img) {
        float r[3];
        position(ent, img, r);
        if (kind == APP_M) {
            const float m = ent >= 0 ? cell[ent].mass : body[4 * ent_body(ent)];
            const float v[4] = {m, r[0], r[1], r[2]};
            M.put(s.mcnt++, v);
        } else {
            float v[HSZ];
            memcpy(v, &rec[(size_t)HSZ * ent], sizeof v);
            v[1] = r[0];
            v[2] = r[1];
            v[3] = r[2];
            if (kind == APP_Q)
                Q.put(s.qcnt++, v);
            else
                H.put(s.hcnt++, v);
        }
        if (dp.log_calls) {
            log_app.push_back(kind);
            log_app.push_back(ent);
            log_app.push_back(img);
        }
    }
    static float sq(float a) { return a * a; }

    // DLRcritMAC, nln/mac.c:1774-1844
    void mac(Sink &s, int32_t sink_ent, const std::vector<Src> &sv, std::vector<int> &result) {
        const size_t app0 = log_app.size() / 3, src0 = log_src.size() / 3;
        result.assign(sv.size(), MAC_SPLIT_SINK);
        stats[0]++;
        for (size_t i = 0; i < sv.size(); i++) {
            const int32_t e = sv[i].ent;
            const int img = sv[i].img;
            // a node entity is a cell (Sub_Flags != 0): single-body leaves are bodies
            const bool sf = e >= 0;
            if (!sf && !s.isbody) {
                result[i] = MAC_SPLIT_SINK;
                continue;
            }
            float r[3];
            position(e, img, r);
            if (e >= 0) {
                const Cell &c = cell[e];
                const float dr2 = sq(r[0] - s.pos[0]) + sq(r[1] - s.pos[1]) + sq(r[2] - s.pos[2]);
                const bool isquad = p2cut && c.daughters >= p2cut;
                const bool ishexa = p4cut && c.daughters >= p4cut;
                const float smallest = ishexa ? c.rcrit_h : (isquad ? c.rcrit_q : c.rcrit);
                if (dr2 > sq(c.rcrit + c.bmax)) {
                    append(s, APP_M, e, img);
                    s.interactions += c.daughters;
                    result[i] = MAC_ACCEPT;
                } else if (isquad && dr2 > sq(c.rcrit_q + c.bmax)) {
                    append(s, APP_Q, e, img);
                    s.interactions += c.daughters;
                    result[i] = MAC_ACCEPT;
                } else if (ishexa && dr2 > sq(c.rcrit_h + c.bmax)) {
                    append(s, APP_H, e, img);
                    s.interactions += c.daughters;
                    result[i] = MAC_ACCEPT;
                } else {
                    result[i] = dp.dlfac * s.bmax > smallest ? MAC_SPLIT_SINK : MAC_SPLIT_SRC;
                }
                // terminal traversal: the source's bodies, one by one
                if (result[i] == MAC_SPLIT_SRC && (!dp.gated || c.daughters <= dp.leaf_max_n)) {
                    result[i] = MAC_ACCEPT;
                    s.interactions += c.daughters;
                    const int64_t b0 = tree.first[e];
                    for (int64_t b = b0; b < b0 + c.daughters; b++) {
                        if (body[4 * b] != body[0])
                            throw std::runtime_error("bad particle mass: the terminal expansion "
                                                     "needs equal masses (nln/mac.c:1824-1829)");
                        append(s, APP_M, body_ent(b), img);
                    }
                }
            } else {
                // body-body
                append(s, APP_M, e, img);
                s.interactions++;
                result[i] = MAC_ACCEPT;
            }
        }
        if (dp.log_calls) {
            for (size_t i = 0; i < sv.size(); i++) {
                log_src.push_back(sv[i].ent);
                log_src.push_back(sv[i].img);
                log_src.push_back(result[i]);
            }
            log_calls.push_back(sink_ent);
            log_calls.push_back((int64_t)src0);
            log_calls.push_back((int64_t)sv.size());
            log_calls.push_back((int64_t)app0);
        }
    }

    // ------------------------------------------------------------ emission
    int64_t emit_blocks(SoA &st, int64_t b0, int64_t b1) {
        const int64_t off = (int64_t)src.size();
        for (int64_t b = b0; b < b1; b++) {
            const float *p = st.block(b);
            src.insert(src.end(), p, p + (size_t)st.sz * NSSE);
        }
        return off;
    }
    void emit_call(int routine, int64_t first, int64_t m, SoA &st, int64_t b0, int64_t b1) {
        const int64_t off = emit_blocks(st, b0, b1);
        const int64_t v[6] = {routine, first, m, off, (b1 - b0) * NSSE, st.sz};
        calls.insert(calls.end(), v, v + 6);
        stats[routine == 0 ? 1 : 2]++;
        stats[routine == 0 ? 9 : 10] += m * (b1 - b0) * NSSE;
    }
    // grav_qns_flush, nln/mac.c:459-475
    void qns_flush() {
        if (!qq.inuse)
            return;
        const int64_t v[6] = {qq.sink_base, qq.m, (int64_t)qseg.size() / 2, (int64_t)src.size(),
                              (int64_t)(qq.source.size() / QSZ), 0};
        qflush.insert(qflush.end(), v, v + 6);
        qseg.insert(qseg.end(), qq.seg.begin(), qq.seg.end());
        src.insert(src.end(), qq.source.begin(), qq.source.end());
        stats[3]++;
        qq.sink_base += qq.m;
        qq.inuse = false;
        qq.m = 0;
        qq.seg.clear();
        qq.source.clear();
    }
    // grav_qns_queue, nln/mac.c:394-428
    void qns_queue(int64_t base_abs, int64_t m, int64_t b0, int64_t b1) {
        if (!qq.inuse) {
            qq.sink_base = base_abs;
            qq.inuse = true;
        }
        const uint32_t len0 = (uint32_t)(qq.source.size() / QSZ);
        if (base_abs < qq.sink_base + qq.m)
            throw std::runtime_error("repeated or out-of-sequence sink in the pQ1 queue");
        while (base_abs > qq.sink_base + qq.m) {  // zero-work placeholders for the gap
            qq.seg.push_back(len0);
            qq.seg.push_back(0);
            qq.m++;
        }
        const uint32_t n = (uint32_t)((b1 - b0) * NSSE);
        for (int64_t b = b0; b < b1; b++) {
            const float *p = Q.block(b);
            qq.source.insert(qq.source.end(), p, p + (size_t)QSZ * NSSE);
        }
        for (int64_t i = 0; i < m; i++) {
            qq.seg.push_back(len0);
            qq.seg.push_back(n);
        }
        qq.m += m;
        stats[10] += m * (int64_t)n;
    }
    // mxn_hexa, nln/mac.c:1180-1272 (the GPU branch: a slot is always granted;
    // with do_pH = 0 the reference runs the same work on the CPU, emitted alike)
    void mxn_hexa(Sink &s) {
        const int64_t n0 = s.hcnt_done / NSSE, n1 = s.hcnt / NSSE;
        if (n1 == n0)
            throw std::runtime_error("mxn_hexa called with n == 0");
        const int64_t mb = (n1 - n0 > dp.hblock) ? 1 : dp.hblock / (n1 - n0);
        for (int64_t p = 0; p < s.daughters; p += mb)
            emit_call(0, s.first + p, std::min(mb, s.daughters - p), H, n0, n1);
        s.hcnt_done = n1 * NSSE;
    }
    // mxn_quad2, nln/mac.c:941-1024 (the GPU branch)
    void mxn_quad2(Sink &s) {
        const int64_t n0 = s.qcnt_done / NSSE, n1 = s.qcnt / NSSE;
        if (dp.do_pQ && n1 > n0) {
            if (dp.do_pQL && s.daughters < 256 && s.first >= qq.sink_base + qq.m) {
                qns_queue(s.first, s.daughters, n0, n1);
                if (qq.m >= 512)
                    qns_flush();
            } else {
                emit_call(1, s.first, s.daughters, Q, n0, n1);
            }
        } else if (n1 > n0) {
            // no GPU: the reference evaluates these on the CPU per body; emitted as a call
            // all the same (its routine says which kernel the reference would use)
            emit_call(1, s.first, s.daughters, Q, n0, n1);
        }
        s.qcnt_done = n1 * NSSE;
    }
    // the leaf: nln/mac.c:506-700 -- each list's remainder, padded to NSSE
    void leaf(Sink &s, int64_t slot) {
        int64_t *t = &tails[6 * slot];
        const int64_t hn = (s.hcnt + NSSE - 1) / NSSE * NSSE;
        H.zero(s.hcnt, hn);
        t[0] = emit_blocks(H, s.hcnt_done / NSSE, hn / NSSE);
        t[1] = hn - s.hcnt_done;
        const int64_t qn = (s.qcnt + NSSE - 1) / NSSE * NSSE;
        Q.zero(s.qcnt, qn);
        t[2] = emit_blocks(Q, s.qcnt_done / NSSE, qn / NSSE);
        t[3] = qn - s.qcnt_done;
        const int64_t mn = (s.mcnt + NSSE - 1) / NSSE * NSSE;
        M.zero(s.mcnt, mn);
        t[4] = emit_blocks(M, 0, mn / NSSE);
        t[5] = mn;
        stats[6] += t[1];
        stats[7] += t[3];
        stats[8] += t[5];
        inter[slot] = s.interactions;
        if (s.interactions != (int64_t)nimage * tree.n)
            stats[11]++;  // the reference aborts here (nln/mac.c:734)
    }

    // ------------------------------------------------------------ walk
    Sink make_sink(int32_t ent) const {
        Sink s{};
        if (ent >= 0) {
            const Cell &c = cell[ent];
            memcpy(s.pos, c.pos, sizeof s.pos);
            s.bmax = c.bmax;
            s.daughters = c.daughters;
            s.isbody = false;
            s.first = tree.first[ent];
        } else {
            const int64_t b = ent_body(ent);
            memcpy(s.pos, &body[4 * b + 1], sizeof s.pos);
            s.bmax = 0.0f;
            s.daughters = 1;
            s.isbody = true;
            s.first = b;
        }
        return s;
    }
    // InheritSinkNlogN from -> to (nln/mac.c:737-846): the M x N work of `from`
    // the first time, then the counts
    Sink inherit(Sink &from, int32_t to_ent) {
        if (!from.processed) {
            if (p4cut && dp.hblock && from.daughters >= dp.min_hsink &&
                from.hcnt - from.hcnt_done >= dp.min_hsrc)
                mxn_hexa(from);
            if (p2cut && dp.hblock && from.daughters >= dp.min_qsink &&
                from.qcnt - from.qcnt_done >= dp.min_qsrc)
                mxn_quad2(from);
            from.processed = true;
        }
        Sink to = make_sink(to_ent);
        to.mcnt = from.mcnt;
        to.qcnt = from.qcnt;
        to.hcnt = from.hcnt;
        to.qcnt_done = from.qcnt_done;
        to.hcnt_done = from.hcnt_done;
        to.interactions = from.interactions;
        return to;
    }
    void walk(Sink &s, int32_t ent, std::vector<Src> pending) {
        std::vector<Src> pass;
        std::vector<int> res;
        while (!pending.empty()) {
            mac(s, ent, pending, res);
            std::vector<Src> opened;
            for (size_t i = 0; i < pending.size(); i++) {
                if (res[i] == MAC_SPLIT_SINK) {
                    pass.push_back(pending[i]);
                } else if (res[i] == MAC_SPLIT_SRC) {
                    const int32_t e = pending[i].ent;
                    for (int64_t k = child_off[e]; k < child_off[e + 1]; k++)
                        opened.push_back({children[k], pending[i].img});
                }
            }
            pending.swap(opened);
        }
        if (s.isbody) {
            leaf(s, ent_body(ent));
            return;
        }
        for (int64_t k = child_off[ent]; k < child_off[ent + 1]; k++) {
            Sink d = inherit(s, children[k]);
            walk(d, children[k], pass);
        }
    }

    void run() {
        const int64_t n = tree.n, nn = tree.nnodes();
        // the image offsets (nln/mac.c:91-111)
        const int f = 2 * nimage_side + 1;
        offset.assign(3 * (size_t)nimage, 0.0f);
        for (int i = -nimage_side; i <= nimage_side; i++)
            for (int j = -nimage_side; j <= nimage_side; j++)
                for (int k = -nimage_side; k <= nimage_side; k++) {
                    const int idx = (i + nimage_side) * f * f + (j + nimage_side) * f + (k + nimage_side);
                    offset[3 * idx + 0] = (float)(i * tp.box);
                    offset[3 * idx + 1] = (float)(j * tp.box);
                    offset[3 * idx + 2] = (float)(k * tp.box);
                }
        body.resize(4 * (size_t)n);
        for (int64_t i = 0; i < n; i++) {
            body[4 * i] = (float)tree.mass[i];
            for (int a = 0; a < 3; a++) body[4 * i + 1 + a] = (float)tree.pos[3 * i + a];
        }
        cell.resize(nn);
        rec.assign((size_t)HSZ * nn, 0.0f);
        for (int64_t c = 0; c < nn; c++) {
            Cell &C = cell[c];
            const double *raw = &tree.raw[35 * (size_t)c];
            C.mass = (float)raw[0];
            for (int a = 0; a < 3; a++) C.pos[a] = (float)tree.cofm[3 * c + a];
            C.bmax = (float)tree.bmax[c];
            C.daughters = tree.count[c];
            if (tp.mac == 1) {
                // radii of orders 0, 2, 4 (a cell too small for an order has none: its
                // test never passes -- isquad / ishexa are false for it anyway)
                const double *rc = &tree.rcrit[3 * (size_t)c];
                C.rcrit = (float)rc[0];
                C.rcrit_q = std::isnan(rc[1]) ? (float)rc[0] : (float)rc[1];
                C.rcrit_h = std::isnan(rc[2]) ? C.rcrit_q : (float)rc[2];
            } else {
                // opening angle: the top order P at bmax / theta (the Bmax MAC,
                // nln/cofm.c:568-569) and each lower order k where its truncation
                // bound (b / r)^(k+1) is the top order's, theta^(P+1)
                const int P = tp.order;
                const double b = tree.bmax[c];
                auto rk = [&](int k) { return (float)(b * pow(tp.theta, -(double)(P + 1) / (k + 1))); };
                C.rcrit = rk(0);
                C.rcrit_q = P >= 2 ? rk(2) : C.rcrit;
                C.rcrit_h = P >= 4 ? rk(4) : C.rcrit_q;
            }
            double r29[HSZ];
            raw_to_record<double>(raw, &tree.cofm[3 * c], tree.bmax[c] > 0 ? tree.bmax[c] : 1.0, r29);
            r29[5] = r29[6] = r29[7] = 0.0;  // no dipole about the centre of mass
            for (int k = 0; k < HSZ; k++) rec[(size_t)HSZ * c + k] = (float)r29[k];
        }
        // daughters of every node as walk entities: child nodes, single-body leaves as
        // bodies, and the bodies of a leaf of several (coincident) bodies
        child_off.assign(nn + 1, 0);
        children.clear();
        for (int64_t c = 0; c < nn; c++) {
            if (tree.child[c] >= 0) {
                for (int k = 0; k < tree.nchild[c]; k++) {
                    const int32_t d = tree.child[c] + k;
                    children.push_back(tree.child[d] < 0 && tree.count[d] == 1 ? body_ent(tree.first[d]) : d);
                }
            } else if (tree.count[c] > 1) {
                for (int64_t b = tree.first[c]; b < tree.first[c] + tree.count[c]; b++)
                    children.push_back(body_ent(b));
            }
            child_off[c + 1] = (int64_t)children.size();
        }
        tails.assign(6 * (size_t)n, 0);
        inter.assign(n, 0);
        // the root: one source per image (WalkInitSrc / WalkInitSrcPeriodic,
        // nln/mac.c:279-291)
        const int32_t root = tree.child[0] < 0 && tree.count[0] == 1 ? body_ent(0) : 0;
        std::vector<Src> src0;
        for (int i = 0; i < nimage; i++) src0.push_back({root, i});
        Sink s = make_sink(root);
        walk(s, root, src0);
        qns_flush();
        stats[4] = (int64_t)src.size();
        stats[5] = (int64_t)calls.size() / 6;
    }
};

}  // namespace
}  // namespace pgrav

using namespace pgrav;

extern "C" {

void pgrav_dl_params_default(pgrav_dl_params *d) {
    if (!d)
        return;
    memset(d, 0, sizeof *d);
    d->dlfac = 6.0f;      // nln/main.c:536
    d->leaf_max_n = 56;   // quad_ncut * 8, nln/main.c:544
    d->gated = 1;
    // mxn_cuda, nln/main.c:215-225
    d->hblock = 16 * 1024 * 1024;
    d->min_qsink = 32;
    d->min_hsink = 32;
    d->min_qsrc = 128;
    d->min_hsrc = 128;
    d->do_pQ = 1;
    d->do_pQL = 1;
    d->do_pH = 1;
    d->log_calls = 0;
}

void *pgrav_dl_walk(const double *pos3, const double *mass, int64_t n, const pgrav_params *p,
                    const pgrav_dl_params *d) {
    clear_error();
    if (!pos3 || !mass || n <= 0 || !p || !d) {
        fail("pgrav_dl_walk: bad arguments");
        return nullptr;
    }
    if (n >= (int64_t)1 << 31) {
        fail("pgrav_dl_walk: at most 2^31 - 1 bodies");
        return nullptr;
    }
    if (!(d->dlfac > 0) || d->leaf_max_n < 1 || d->hblock < 0 || d->min_hsink < 0 ||
        d->min_qsink < 0 || d->min_hsrc < 0 || d->min_qsrc < 0) {
        fail("pgrav_dl_walk: dlfac > 0, leaf_max_n >= 1 and non-negative M x N limits");
        return nullptr;
    }
    if (p->periodic == 1) {
        fail("pgrav_dl_walk: periodic 1 (nearest image + Ewald table) has no DL walk; use 0 or 2");
        return nullptr;
    }
    DLWalk *W = new (std::nothrow) DLWalk();
    if (!W) {
        fail("out of host memory");
        return nullptr;
    }
    const std::string bad = tree_params_from(p, W->tp);
    if (!bad.empty()) {
        fail("pgrav_dl_walk: %s", bad.c_str());
        delete W;
        return nullptr;
    }
    W->dp = *d;
    W->tp.leaf_max = 1;  // the reference's tree: a body per leaf (coincident ones aside)
    W->p4cut = W->tp.order >= 4 ? W->tp.hexa_ncut : 0;
    W->p2cut = W->tp.order >= 2 ? W->tp.quad_ncut : 0;
    W->nimage_side = W->tp.periodic == 2 ? W->tp.nimage : 0;
    const int f = 2 * W->nimage_side + 1;
    W->nimage = f * f * f;
    try {
        build_tree(pos3, mass, n, W->tp, W->tree);
        W->run();
    } catch (const std::exception &e) {
        fail("pgrav_dl_walk: %s", e.what());
        delete W;
        return nullptr;
    }
    return W;
}

void pgrav_dl_destroy(void *h) { delete (DLWalk *)h; }

int pgrav_dl_stats(void *h, int64_t *out16) {
    DLWalk *W = (DLWalk *)h;
    if (!W || !out16) {
        fail("pgrav_dl_stats: null argument");
        return -1;
    }
    memcpy(out16, W->stats, sizeof W->stats);
    out16[12] = W->tree.n;
    out16[13] = W->tree.nnodes();
    out16[14] = W->nimage;
    out16[15] = (int64_t)W->log_calls.size() / 4;
    return 0;
}

int64_t pgrav_dl_array(void *h, int which, void *out) {
    DLWalk *W = (DLWalk *)h;
    if (!W) {
        fail("pgrav_dl_array: null handle");
        return -1;
    }
    const void *p = nullptr;
    size_t bytes = 0;
    int64_t count = 0;
    auto take = [&](const auto &v, int per) {
        p = v.data();
        bytes = v.size() * sizeof(v[0]);
        count = (int64_t)(v.size() / per);
    };
    std::vector<float> cells;
    std::vector<int64_t> perm;
    switch (which) {
    case PGRAV_DL_CELLS:
        cells.resize(8 * W->cell.size());
        for (size_t c = 0; c < W->cell.size(); c++) {
            const Cell &C = W->cell[c];
            const float v[8] = {C.mass, C.pos[0], C.pos[1], C.pos[2], C.rcrit, C.bmax, C.rcrit_q, C.rcrit_h};
            memcpy(&cells[8 * c], v, sizeof v);
        }
        take(cells, 8);
        break;
    case PGRAV_DL_RECORDS: take(W->rec, 29); break;
    case PGRAV_DL_BODIES: take(W->body, 4); break;
    case PGRAV_DL_PERM: take(W->tree.perm, 1); break;
    case PGRAV_DL_CHILD_OFF: take(W->child_off, 1); break;
    case PGRAV_DL_CHILDREN: take(W->children, 1); break;
    case PGRAV_DL_CALLS: take(W->calls, 6); break;
    case PGRAV_DL_SOURCES: take(W->src, 1); break;
    case PGRAV_DL_QFLUSH: take(W->qflush, 6); break;
    case PGRAV_DL_QSEG: take(W->qseg, 2); break;
    case PGRAV_DL_TAILS: take(W->tails, 6); break;
    case PGRAV_DL_INTERACTIONS: take(W->inter, 1); break;
    case PGRAV_DL_LOG_CALLS: take(W->log_calls, 4); break;
    case PGRAV_DL_LOG_SOURCES: take(W->log_src, 3); break;
    case PGRAV_DL_LOG_APPENDS: take(W->log_app, 3); break;
    case PGRAV_DL_NODE_INFO:
        perm.resize(4 * (size_t)W->tree.nnodes());
        for (int64_t c = 0; c < W->tree.nnodes(); c++) {
            perm[4 * c + 0] = W->tree.first[c];
            perm[4 * c + 1] = W->tree.count[c];
            perm[4 * c + 2] = W->tree.child[c] >= 0 ? 1 : 0;
            perm[4 * c + 3] = W->tree.level[c];
        }
        take(perm, 4);
        break;
    default:
        fail("pgrav_dl_array: unknown array %d", which);
        return -1;
    }
    if (out && bytes)
        memcpy(out, p, bytes);
    return count;
}

}  // extern "C"
