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
//  - subtract_background = 1 restates DLRcritMACsb (nln/mac.c:1565-1771) and the
//    parts of the walk it brings: records of the moments less each cell's
//    uniform cube of the mean density rho0 (mpole_add_cube, nln/cofm.c:626-648,
//    applied at :348 with the cell's key-cube edge); the sink's near cube
//    (clevel / cen / cr / cr2 set on inheritance, nln/mac.c:747-748, :824-843);
//    anti-masses for the background of split cells, empty octants and bodies
//    outside the near cube; a background flag per source, carried by the
//    source's daughters when it is split (the walk's flags word, as the image
//    index is); MMvec segments for small central-image cells, batched per sink
//    by mxn_mono (nln/mac.c:877-936) into the grav_mns_queue / grav_mns_flush
//    queue (nln/mac.c:352-457); and, at the leaf, -rho0 times the near cube's
//    field (nln/mac.c:688-694), here from the closed form of a uniform cube's
//    attraction in FP64.
//
// Everything here is host code; the GPU evaluates the emitted calls through the
// drop-in ABI.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <memory>
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
// list appends, and (subtract_background) their variants: the codes of the call log
enum { APP_M = 0, APP_Q = 1, APP_H = 2, APP_QB = 3, APP_HB = 4, APP_ANTI_BODY = 5,
       APP_ANTI_CELL = 6, APP_ANTI_EMPTY = 8, APP_MM = 20 };
const int CHUBITS = 42;               // (KEYBITS - 1) / NDIM of the two-word key: "no clevel"
const int MNSQ_MAX = 32768 + 1024;    // nln/mac.c:157-158
const int MNSQ_MAX_M = 131072;

// A walk entity: a tree node (>= 0) or a body, -(sorted slot + 1).
inline int32_t body_ent(int64_t slot) { return (int32_t)(-(slot + 1)); }
inline int64_t ent_body(int32_t e) { return -(int64_t)e - 1; }

struct Src {
    int32_t ent;
    int32_t img;  // offset index (nln/mac.c:91-111 numbering)
    uint8_t bs;   // BACKGROUND_FLAG: this source's background is already accounted for
};

// The field at r of a cube [-a, a]^3 of unit density (unit G): sum over the faces
// x = +-a of the double integral of 1/|u| over the face in the coordinates
// u = source point - r, by the prism formula (Nagy 1966): F(c, y, z) = y ln(z + R) +
// z ln(y + R) - c atan(y z / (c R)), R = |(c, y, z)|, is the antiderivative of
// 1/R in y and z.  Logarithms of near-cancelling sums are taken in the
// equivalent form ln((c^2 + y^2) / (R - z)); terms whose factor is zero are zero.
double cube_face(double c, double y, double z) {
    const double R = sqrt(c * c + y * y + z * z);
    auto ylog = [&](double p, double q) {  // p ln(q + R)
        if (p == 0.0)
            return 0.0;
        const double t = q >= 0 ? q + R : (c * c + p * p) / (R - q);
        return p * log(t);
    };
    double f = ylog(y, z) + ylog(z, y);
    if (c != 0.0)
        f -= c * atan(y * z / (c * R));
    return f;
}

void cube_field(const double r[3], double a, double g[3]) {
    for (int ax = 0; ax < 3; ax++) {
        const int b = (ax + 1) % 3, c = (ax + 2) % 3;
        const double lo[3] = {-a - r[0], -a - r[1], -a - r[2]};
        const double hi[3] = {a - r[0], a - r[1], a - r[2]};
        double s = 0.0;
        for (int face = 0; face < 2; face++) {
            const double u = face ? hi[ax] : lo[ax];
            const double sf = face ? -1.0 : 1.0;  // int u / |u|^3 du = [-1/|u|]_lo^hi
            s += sf * (cube_face(u, hi[b], hi[c]) - cube_face(u, hi[b], lo[c]) -
                       cube_face(u, lo[b], hi[c]) + cube_face(u, lo[b], lo[c]));
        }
        g[ax] = s;
    }
}

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
    // subtract_background
    bool done;
    int clevel;  // the near cube's level (CHUBITS: none yet)
    float cen[3], cr, cr2;
    int near;
    float fmass;
    int64_t mmcnt, mmcnt_done, mmterms, mmterms_done;
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

// The emitted source blocks, the walk's largest output, in fixed chunks of 64 MB
// that are filled in order: it grows without reallocating, where a doubling
// vector would copy itself at every growth and touch about twice its final size
// of fresh memory.  Offsets are global, as in one array.
struct SourceStore {
    static constexpr size_t CH = (size_t)1 << 24;  // floats per chunk
    std::vector<std::unique_ptr<float[]>> chunk;
    size_t n = 0;
    size_t size() const { return n; }
    void append(const float *p, size_t k) {
        while (k) {
            if (n == chunk.size() * CH)
                chunk.emplace_back(new float[CH]);
            const size_t at = n % CH, t = std::min(CH - at, k);
            memcpy(chunk.back().get() + at, p, t * sizeof(float));
            p += t;
            k -= t;
            n += t;
        }
    }
    float &operator[](size_t i) { return chunk[i / CH][i % CH]; }
    void copy_out(float *dst) const {
        for (size_t c = 0, left = n; left; c++) {
            const size_t t = std::min(CH, left);
            memcpy(dst + c * CH, chunk[c].get(), t * sizeof(float));
            left -= t;
        }
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
    SourceStore src;              // every emitted source block
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

    // ---- subtract_background
    bool sb = false;
    double rho0 = 0.0;  // m0 / (2 r0)^3 (nln/cofm.c:62)
    float r0 = 0.0f, nx = 0.0f;
    struct UCell {
        float halfsz, mass;
    } ucell[CHUBITS];
    std::vector<float> geom;      // 5 per node: corner x, y, z, edge, octant mask
    std::vector<uint32_t> icorner;  // 3 per node: the key cube's integer corner at its level
    std::vector<float> ubs;       // 4 per node: umass, uxxxx, uyyyy, uxxyy
    std::vector<int64_t> body_leaf;
    std::vector<uint32_t> mm;     // 2 per MMvec entry (the sink stack): base, length
    std::vector<uint32_t> mseg;   // 2 per emitted segment
    std::vector<int64_t> mflush;  // 8 per pMM_mnss_sK1 flush
    std::vector<uint16_t> mss_index;
    std::vector<uint32_t> mss_seg;
    std::vector<int32_t> mss_source_n;
    std::vector<int64_t> mcalls;   // 6 per host mxn_mono run
    std::vector<int64_t> mmtails;  // 3 per body slot
    std::vector<double> background;  // 3 per body slot
    std::vector<float> sbsink;       // 6 per body slot
    std::vector<float> log_sb;       // 8 per MAC call
    std::vector<int32_t> log_flags;  // 2 per logged source
    int64_t sbstats[8] = {0};
    // tail_spans: per list, the source offset of each stack block's latest
    // snapshot on the current sink path (-1: none), the leaves' runs of them
    std::vector<int64_t> blk_off[3];
    std::vector<int64_t> tail_runs;   // 3 per run: list, source offset, records
    std::vector<int64_t> tail_index;  // 2 per body slot: first run, runs
    // the mns queue (nln/mac.c:160-173)
    struct {
        bool inuse = false;
        int64_t sink_base = 0, m = 0;
        std::vector<uint16_t> ss_index;
        std::vector<uint32_t> ss_seg;  // 2 per entry
        std::vector<int32_t> source_n;
        std::vector<uint32_t> seg;     // 2 per segment
    } mq;

    // ------------------------------------------------------------ MAC
    void position(int32_t ent, int img, float r[3]) const {
        const float *p = ent >= 0 ? cell[ent].pos : &body[4 * ent_body(ent) + 1];
        const float *o = &offset[3 * img];
        r[0] = p[0] + o[0];
        r[1] = p[1] + o[1];
        r[2] = p[2] + o[2];
    }
    void log_append(int code, int32_t ent, int img) {
        if (dp.log_calls) {
            log_app.push_back(code);
            log_app.push_back(ent);
            log_app.push_back(img);
        }
    }
    // Returns the mass written.  Codes APP_ANTI_* write `mass` at the source's
    // position plus offset; APP_QB / APP_HB the background-flagged records.
    float append(Sink &s, int kind, int32_t ent, int img, int code = -1, float mass = 0.0f) {
        float r[3];
        position(ent, img, r);
        float m;
        if (kind == APP_M) {
            m = code >= APP_ANTI_BODY ? mass : (ent >= 0 ? cell[ent].mass : body[4 * ent_body(ent)]);
            const float v[4] = {m, r[0], r[1], r[2]};
            M.put(s.mcnt++, v);
        } else {
            float v[HSZ];
            memcpy(v, &rec[(size_t)HSZ * ent], sizeof v);
            v[1] = r[0];
            v[2] = r[1];
            v[3] = r[2];
            if (code == APP_QB) {
                v[0] = v[0] + ucell[node_level(ent)].mass;  // appendQdvec, nln/mac.c:1444
            } else if (code == APP_HB) {                    // appendHdvec, nln/mac.c:1470-1476
                const float *u = &ubs[4 * (size_t)ent];
                v[0] = u[0];
                v[20] = u[1];
                v[24] = u[2];
                v[22] = u[3];
            }
            m = v[0];
            if (kind == APP_Q)
                Q.put(s.qcnt++, v);
            else
                H.put(s.hcnt++, v);
        }
        log_append(code >= 0 ? code : kind, ent, img);
        return m;
    }
    int node_level(int32_t c) const { return tree.level[c]; }
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

    // DLRcritMACsb, nln/mac.c:1565-1771.  FP32 as the reference's locals; the
    // background flags of `sv` are updated in place (they travel with the sources).
    static float d2(const float *a, const float *b) {
        return sq(a[0] - b[0]) + sq(a[1] - b[1]) + sq(a[2] - b[2]);
    }
    // outside the sink's near cube along some axis (the anti-mass test, nln/mac.c:1671-1673)
    static bool outside(const Sink &s, const float m[3]) {
        return fabsf(m[0] - s.cen[0]) > s.cr || fabsf(m[1] - s.cen[1]) > s.cr ||
               fabsf(m[2] - s.cen[2]) > s.cr;
    }
    // the centre of a key cube plus offset, as CellCorner + VV(+= 0.5 cellsz + offset)
    void cube_centre(const float corner[3], float edge, int img, float m[3]) const {
        const float *o = &offset[3 * img];
        for (int a = 0; a < 3; a++) m[a] = corner[a] + (0.5f * edge + o[a]);
    }
    std::vector<uint8_t> bs_in;  // mac_sb's incoming flags (for the log)
    void mac_sb(Sink &s, int32_t sink_ent, std::vector<Src> &sv, std::vector<int> &result) {
        const size_t app0 = log_app.size() / 3, src0 = log_src.size() / 3;
        result.assign(sv.size(), MAC_SPLIT_SINK);
        bs_in.resize(sv.size());
        for (size_t i = 0; i < sv.size(); i++) bs_in[i] = sv[i].bs;
        stats[0]++;
        const int central = nimage >> 1;
        if (!s.done) {
            for (size_t i = 0; i < sv.size(); i++) {
                const int32_t e = sv[i].ent;
                const int img = sv[i].img;
                bool bs = sv[i].bs != 0;
                float r[3];
                position(e, img, r);
                if (e >= 0) {
                    const Cell &c = cell[e];
                    const int lev = tree.level[e];
                    const bool isquad = p2cut && c.daughters >= p2cut;
                    const bool ishexa = p4cut && c.daughters >= p4cut;
                    const float smallest = ishexa ? c.rcrit_h : (isquad ? c.rcrit_q : c.rcrit);
                    float dr2, bmax;
                    if (s.clevel != CHUBITS && lev < s.clevel) {
                        dr2 = d2(r, s.cen);
                        bmax = 1.2f * s.cr;
                    } else {
                        dr2 = d2(r, s.pos);
                        bmax = s.bmax;
                    }
                    if (!bs && lev == s.clevel && d2(r, s.cen) < s.cr2) {
                        bs = true;
                        sv[i].bs = 1;
                        s.near++;
                    }
                    if (isquad && dr2 > sq(c.rcrit_q + bmax)) {
                        s.fmass += append(s, APP_Q, e, img, bs ? APP_QB : APP_Q);
                        s.interactions += c.daughters;
                        result[i] = MAC_ACCEPT;
                    } else if (ishexa && dr2 > sq(c.rcrit_h + bmax)) {
                        s.fmass += append(s, APP_H, e, img, bs ? APP_HB : APP_H);
                        s.interactions += c.daughters;
                        result[i] = MAC_ACCEPT;
                    } else {
                        result[i] = (dp.dlfac * s.bmax > smallest && s.daughters > dp.leaf_max_n)
                                        ? MAC_SPLIT_SINK
                                        : MAC_SPLIT_SRC;
                    }
                } else {
                    // body-body, and the anti-mass of the body's leaf cube unless inside
                    // the near cube (nln/mac.c:1658-1686)
                    s.fmass += append(s, APP_M, e, img);
                    s.interactions++;
                    result[i] = MAC_ACCEPT;
                    if (!bs) {
                        const float *g = &geom[5 * (size_t)body_leaf[ent_body(e)]];
                        float mx[3];
                        cube_centre(g, g[3], img, mx);
                        if (outside(s, mx)) {
                            const double c3 = g[3];
                            const float am = (float)(-rho0 * c3 * c3 * c3);
                            s.fmass += append(s, APP_M, e, img, APP_ANTI_BODY, am);
                            sbstats[3]++;
                        } else {
                            s.near++;
                        }
                    }
                }
                if (result[i] == MAC_SPLIT_SRC && !bs) {
                    const int lev = tree.level[e];
                    const float *g = &geom[5 * (size_t)e];
                    const int sf = (int)g[4];
                    if (lev >= s.clevel) {
                        // the whole cell's background, once: its daughters carry the flag
                        s.fmass += append(s, APP_M, e, img, APP_ANTI_CELL, -ucell[lev].mass);
                        sv[i].bs = 1;
                        sbstats[4]++;
                    } else if (sf != 0xff && lev > 4) {
                        // empty octants (nln/mac.c:1707-1734), in key-digit order
                        const uint32_t *ic = &icorner[3 * (size_t)e];
                        const float factor = (float)tree.size / (float)(1u << (lev + 1));
                        for (int j = 0; j < 8; j++) {
                            if (sf >> j & 1)
                                continue;
                            float corner[3];
                            for (int a = 0; a < 3; a++)
                                corner[a] = (float)tree.x0[a] +
                                            factor * (float)(2 * ic[a] + (uint32_t)(j >> a & 1));
                            float mx[3];
                            cube_centre(corner, factor, img, mx);
                            if (outside(s, mx)) {
                                s.fmass += append(s, APP_M, e, img, APP_ANTI_EMPTY + j,
                                                  -ucell[lev + 1].mass);
                                sbstats[5]++;
                            } else {
                                s.near++;
                            }
                        }
                    }
                }
                // terminal traversal of a small cell (nln/mac.c:1736-1753)
                if (e >= 0 && result[i] != MAC_ACCEPT && cell[e].daughters <= dp.leaf_max_n) {
                    const int64_t nb = cell[e].daughters, b0 = tree.first[e];
                    if (dp.do_pM && img == central) {
                        if (mm.size() < 2 * (size_t)(s.mmcnt + 1))
                            mm.resize(std::max(2 * (size_t)(s.mmcnt + 1), 2 * mm.size()));
                        sbstats[2]++;
                        mm[2 * s.mmcnt] = (uint32_t)b0;
                        mm[2 * s.mmcnt + 1] = (uint32_t)nb;
                        s.mmcnt++;
                        s.mmterms += nb;
                        log_append(APP_MM, e, img);
                    } else {
                        for (int64_t b = b0; b < b0 + nb; b++) append(s, APP_M, body_ent(b), img);
                    }
                    result[i] = MAC_ACCEPT;
                    s.interactions += nb;
                }
            }
        }
        if (dp.log_calls) {
            for (size_t i = 0; i < sv.size(); i++) {
                log_src.push_back(sv[i].ent);
                log_src.push_back(sv[i].img);
                log_src.push_back(result[i]);
                log_flags.push_back(bs_in[i]);
                log_flags.push_back(sv[i].bs);
            }
            log_calls.push_back(sink_ent);
            log_calls.push_back((int64_t)src0);
            log_calls.push_back((int64_t)sv.size());
            log_calls.push_back((int64_t)app0);
        }
    }
    void log_sink_sb(const Sink &s) {
        if (!dp.log_calls)
            return;
        const float v[8] = {s.cen[0], s.cen[1], s.cen[2], s.cr, s.cr2, (float)s.clevel,
                            s.done ? 1.0f : 0.0f, (float)s.near};
        log_sb.insert(log_sb.end(), v, v + 8);
    }

    // ------------------------------------------------------------ emission
    int64_t emit_blocks(SoA &st, int64_t b0, int64_t b1) {
        const int64_t off = (int64_t)src.size();
        for (int64_t b = b0; b < b1; b++) {
            const float *p = st.block(b);
            src.append(p, (size_t)st.sz * NSSE);
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
        src.append(qq.source.data(), qq.source.size());
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
    // grav_mns_flush, nln/mac.c:430-457: one grav_mnss_CUDA("pMM_mnss_sK1") call
    // (source_base 0: the segments index the body table)
    void mns_flush() {
        if (!mq.inuse)
            return;
        const int64_t v[8] = {mq.sink_base, mq.m, (int64_t)mss_index.size(),
                              (int64_t)mss_source_n.size(), (int64_t)mq.source_n.size(),
                              (int64_t)mseg.size() / 2, (int64_t)mq.seg.size() / 2, 0};
        mflush.insert(mflush.end(), v, v + 8);
        mss_index.insert(mss_index.end(), mq.ss_index.begin(), mq.ss_index.begin() + mq.m);
        mss_seg.insert(mss_seg.end(), mq.ss_seg.begin(), mq.ss_seg.end());
        mss_source_n.insert(mss_source_n.end(), mq.source_n.begin(), mq.source_n.end());
        mseg.insert(mseg.end(), mq.seg.begin(), mq.seg.end());
        sbstats[0]++;
        mq.sink_base += mq.m;
        mq.inuse = false;
        mq.m = 0;
        mq.ss_index.clear();
        mq.ss_seg.clear();
        mq.source_n.clear();
        mq.seg.clear();
    }
    // grav_mns_queue, nln/mac.c:351-392: sinks base .. base + m take the segment run
    // `seg` (seg_n segments, source_n bodies); a gap before them becomes one
    // zero-work entry
    void mns_queue(int64_t base_abs, int64_t m, const uint32_t *seg, int64_t seg_n,
                   int64_t source_n) {
        if (!mq.inuse) {
            mq.sink_base = base_abs;
            mq.inuse = true;
        }
        const int64_t base = base_abs - mq.sink_base;
        if (base_abs < mq.sink_base + mq.m)
            throw std::runtime_error("repeated or out-of-sequence sink in grav_mns_queue");
        auto push_entry = [&](uint32_t b, uint32_t len, int32_t sn) {
            mq.ss_seg.push_back(b);
            mq.ss_seg.push_back(len);
            mq.source_n.push_back(sn);
        };
        if (base_abs > mq.sink_base + mq.m) {
            const int64_t gap = base_abs - (mq.sink_base + mq.m);
            const uint16_t k = (uint16_t)mq.source_n.size();
            push_entry((uint32_t)(mq.seg.size() / 2), 0, 0);
            for (int64_t i = 0; i < gap; i++) mq.ss_index.push_back(k);
            mq.m += gap;
        }
        if (base + m >= MNSQ_MAX_M)
            throw std::runtime_error("mnsq.ss_index overflow");
        const uint16_t k = (uint16_t)mq.source_n.size();
        push_entry((uint32_t)(mq.seg.size() / 2), (uint32_t)seg_n, (int32_t)source_n);
        mq.seg.insert(mq.seg.end(), seg, seg + 2 * seg_n);
        for (int64_t i = 0; i < m; i++) mq.ss_index.push_back(k);
        mq.m += m;
        sbstats[6] += m * source_n;
        if ((int64_t)mq.seg.size() / 2 >= MNSQ_MAX)
            throw std::runtime_error("mnsq.seg_n overflow");
        if ((int64_t)mq.source_n.size() >= MNSQ_MAX / 16)
            throw std::runtime_error("mnsq.ss_len overflow");
    }
    // mxn_mono, nln/mac.c:877-936: the sink's MMvec segments since the last run,
    // queued for the GPU when the sink starts past the queue, else run on the host
    // for each body (recorded as a call)
    void mxn_mono(Sink &s) {
        const int64_t seg_n = s.mmcnt - s.mmcnt_done, source_n = s.mmterms - s.mmterms_done;
        if (dp.do_pM && s.first >= mq.sink_base + mq.m) {
            // the reference passes the stack's start (MMvec); the walk never queues a
            // sink whose ancestors ran mxn_mono, so mmcnt_done is 0 here
            if (s.mmcnt_done != 0)
                throw std::runtime_error("mxn_mono: queue with segments already done");
            mns_queue(s.first, s.daughters, mm.data(), seg_n, source_n);
            if ((int64_t)mq.seg.size() / 2 >= 32768)
                mns_flush();
        } else {
            const int64_t v[6] = {s.first, s.daughters, (int64_t)mseg.size() / 2, seg_n,
                                  source_n, 0};
            mcalls.insert(mcalls.end(), v, v + 6);
            mseg.insert(mseg.end(), mm.begin() + 2 * s.mmcnt_done, mm.begin() + 2 * s.mmcnt);
            sbstats[1]++;
        }
        s.mmcnt_done = s.mmcnt;
        s.mmterms_done = s.mmterms;
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
    // tail_spans: copy stack blocks [c0 / NSSE, ceil(c1 / NSSE)) of list `kind`
    // once, records past c1 zeroed (they belong to no sink on this path), and
    // point the path's block table at the copies; `undo` restores it on return
    void snapshot(int kind, SoA &st, int64_t c0, int64_t c1,
                  std::vector<std::pair<int64_t, int64_t>> &undo) {
        if (c1 <= c0)
            return;
        const int64_t b0 = c0 / NSSE, b1 = (c1 + NSSE - 1) / NSSE;
        std::vector<int64_t> &bo = blk_off[kind];
        if ((int64_t)bo.size() < b1)
            bo.resize(std::max<int64_t>(b1, 2 * (int64_t)bo.size()), -1);
        for (int64_t b = b0; b < b1; b++) {
            undo.push_back({kind * ((int64_t)1 << 56) + b, bo[b]});
            const int64_t off = (int64_t)src.size();
            const float *p = st.block(b);
            src.append(p, (size_t)st.sz * NSSE);
            if (b == b1 - 1 && c1 % NSSE)
                for (int f = 0; f < st.sz; f++)
                    for (int r = (int)(c1 % NSSE); r < NSSE; r++) src[off + f * NSSE + r] = 0.0f;
            bo[b] = off;
        }
    }
    void undo_snapshots(const std::vector<std::pair<int64_t, int64_t>> &undo) {
        for (size_t k = undo.size(); k-- > 0;) {
            const int kind = (int)(undo[k].first >> 56);
            blk_off[kind][undo[k].first & (((int64_t)1 << 56) - 1)] = undo[k].second;
        }
    }
    // a leaf's remainder of list `kind`, records [c0, c1) padded to NSSE, as runs of
    // snapshot blocks
    void tail_from_snapshots(int kind, int sz, int64_t c0, int64_t c1) {
        const int64_t b0 = c0 / NSSE, b1 = (c1 + NSSE - 1) / NSSE;
        for (int64_t b = b0; b < b1; b++) {
            const int64_t off = blk_off[kind][b];
            if (off < 0)
                throw std::runtime_error("leaf remainder block without a snapshot");
            const size_t nr = tail_runs.size();
            if (b > b0 && tail_runs[nr - 3] == kind &&
                tail_runs[nr - 2] + tail_runs[nr - 1] / NSSE * sz * NSSE == off) {
                tail_runs[nr - 1] += NSSE;
            } else {
                const int64_t v[3] = {kind, off, NSSE};
                tail_runs.insert(tail_runs.end(), v, v + 3);
            }
        }
    }

    // the leaf: nln/mac.c:506-700 -- each list's remainder, padded to NSSE
    void leaf(Sink &s, int64_t slot) {
        if (sb) {
            if (slot >= tree.n - 1)
                mns_flush();  // at the last body (nln/mac.c:505-506)
            int64_t *mt = &mmtails[3 * slot];
            mt[0] = (int64_t)mseg.size() / 2;
            mt[1] = s.mmcnt - s.mmcnt_done;
            mt[2] = s.mmterms - s.mmterms_done;
            mseg.insert(mseg.end(), mm.begin() + 2 * s.mmcnt_done, mm.begin() + 2 * s.mmcnt);
            // the near cube's background, -rho0 times a uniform cube's field
            // (nln/mac.c:688-694)
            double r[3], g[3];
            for (int a = 0; a < 3; a++) r[a] = (double)(s.pos[a] - s.cen[a]);
            cube_field(r, (double)s.cr, g);
            for (int a = 0; a < 3; a++) background[3 * slot + a] = -rho0 * g[a];
            float *q = &sbsink[6 * slot];
            q[0] = s.cen[0];
            q[1] = s.cen[1];
            q[2] = s.cen[2];
            q[3] = s.cr;
            q[4] = s.fmass;
            q[5] = (float)s.near;
        }
        int64_t *t = &tails[6 * slot];
        const int64_t hn = (s.hcnt + NSSE - 1) / NSSE * NSSE;
        const int64_t qn = (s.qcnt + NSSE - 1) / NSSE * NSSE;
        const int64_t mn = (s.mcnt + NSSE - 1) / NSSE * NSSE;
        t[1] = hn - s.hcnt_done;
        t[3] = qn - s.qcnt_done;
        t[5] = mn;
        if (dp.tail_spans) {
            tail_index[2 * slot] = (int64_t)tail_runs.size() / 3;
            tail_from_snapshots(0, HSZ, s.hcnt_done, s.hcnt);
            tail_from_snapshots(1, QSZ, s.qcnt_done, s.qcnt);
            tail_from_snapshots(2, MSZ, 0, s.mcnt);
            tail_index[2 * slot + 1] = (int64_t)tail_runs.size() / 3 - tail_index[2 * slot];
            t[0] = t[2] = t[4] = -1;
        } else {
            H.zero(s.hcnt, hn);
            t[0] = emit_blocks(H, s.hcnt_done / NSSE, hn / NSSE);
            Q.zero(s.qcnt, qn);
            t[2] = emit_blocks(Q, s.qcnt_done / NSSE, qn / NSSE);
            M.zero(s.mcnt, mn);
            t[4] = emit_blocks(M, 0, mn / NSSE);
        }
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
        s.clevel = CHUBITS;
        return s;
    }

    // Each node's key cube (CellCorner arithmetic, nln/physics_generic.c:237-258:
    // corner = Rmin + (Rsize / 2^level) * integer corner, in float; Rmin = the box
    // origin, Rsize = the box) and its octants holding bodies.
    void setup_geom() {
        const int64_t nn = tree.nnodes();
        // integer corners from the first body's key coordinates (the tree's keys)
        const double scale = (double)(1 << 21) / tree.size;
        auto icoord = [&](int64_t b, int a) {
            int64_t q = (int64_t)((tree.pos[3 * b + a] - tree.x0[a]) * scale);
            return (uint32_t)std::max<int64_t>(0, std::min<int64_t>((1 << 21) - 1, q));
        };
        geom.assign(5 * (size_t)nn, 0.0f);
        icorner.assign(3 * (size_t)nn, 0);
        for (int64_t c = 0; c < nn; c++) {
            const int L = tree.level[c];
            const float factor = (float)tree.size / (float)(1u << L);
            for (int a = 0; a < 3; a++) {
                const uint32_t ic = icoord(tree.first[c], a) >> (21 - L);
                icorner[3 * c + a] = ic;
                geom[5 * c + a] = (float)tree.x0[a] + factor * (float)ic;
            }
            geom[5 * c + 3] = factor;
            int sf = 0;
            if (tree.child[c] >= 0) {
                for (int k = 0; k < tree.nchild[c]; k++) {
                    const int32_t d = tree.child[c] + k;
                    int j = 0;
                    for (int a = 0; a < 3; a++) j |= (int)(icoord(tree.first[d], a) >> (21 - L - 1) & 1) << a;
                    sf |= 1 << j;
                }
            } else {
                sf = 0xff;  // a leaf (one body, or coincident ones): no empty octants
            }
            geom[5 * c + 4] = (float)sf;
        }
    }

    // subtract_background: the mean density and the ucell table (SetupCofm,
    // nln/cofm.c:62; SetupGrav, nln/mac.c:237-250, FP32 as there)
    void setup_sb() {
        const int64_t n = tree.n, nn = tree.nnodes();
        double m0 = 0.0;
        for (int64_t i = 0; i < n; i++) m0 += tree.mass[i];
        r0 = (float)(0.5 * tp.box);
        rho0 = m0 / ((2.0 * r0) * (2.0 * r0) * (2.0 * r0));
        nx = dp.nx > 0.0f ? dp.nx : (float)cbrt((double)n);
        float a1 = r0;
        const float e2 = 2.0f * a1;
        float a3 = (float)(rho0 * (double)(e2 * e2 * e2));
        for (int i = 0; i < CHUBITS; i++) {
            ucell[i].halfsz = a1;
            ucell[i].mass = a3;
            a1 *= 0.5f;
            a3 *= 0.125f;
        }
        body_leaf.assign(n, 0);
        mmtails.assign(3 * (size_t)n, 0);
        background.assign(3 * (size_t)n, 0.0);
        sbsink.assign(6 * (size_t)n, 0.0f);
        ubs.assign(4 * (size_t)nn, 0.0f);
    }

    // The expansion basis of every cell: its centre, the raw moments about it, bmax
    // and the critical radii.  By default the tree's (centre of mass); with
    // geometric_center the key cube's centre (nln/cofm.c:141-146: moments about it,
    // the dipole no longer zero) and bmax by CofmFromDaugh's recursion about it
    // (nln/cofm.c:183-222: daughters' distance plus bmax, bodies' distance, capped
    // by the cube's half-diagonal).  The error-tolerance radii are taken about the
    // basis centre; with subtract_background, cells of more than 8 bodies bound the
    // odd terms by the sums of |third| and |fifth| moments with B4 = B6 = 0
    // (nln/cofm.c:356-360, :425-434).
    std::vector<double> bcen, braw, bbmax, brcrit;
    void cell_basis() {
        const int64_t nn = tree.nnodes();
        const bool geo = dp.geometric_center != 0;
        bcen = tree.cofm;
        braw = tree.raw;
        bbmax = tree.bmax;
        brcrit = tree.rcrit;
        if (geo) {
            for (int64_t c = 0; c < nn; c++) {
                const float *g = &geom[5 * (size_t)c];
                for (int a = 0; a < 3; a++) bcen[3 * c + a] = (double)(g[a] + 0.5f * g[3]);
                double *u = &braw[35 * (size_t)c];
                std::fill(u, u + 35, 0.0);
                const double *cc = &bcen[3 * c];
                for (int64_t i = tree.first[c]; i < tree.first[c] + tree.count[c]; i++) {
                    const double d[3] = {tree.pos[3 * i] - cc[0], tree.pos[3 * i + 1] - cc[1],
                                         tree.pos[3 * i + 2] - cc[2]};
                    for (int a = 0; a <= 4; a++)
                        for (int b = 0; a + b <= 4; b++)
                            for (int e = 0; a + b + e <= 4; e++)
                                u[raw_index(a, b, e)] += tree.mass[i] * pow(d[0], a) * pow(d[1], b) * pow(d[2], e);
                }
            }
            // children follow their parents in the node order: deepest first backwards
            for (int64_t c = nn - 1; c >= 0; c--) {
                const double *cc = &bcen[3 * c];
                double bm = 0.0;
                auto dist = [&](const double *p) {
                    const double dx = p[0] - cc[0], dy = p[1] - cc[1], dz = p[2] - cc[2];
                    return sqrt(dx * dx + dy * dy + dz * dz);
                };
                if (tree.child[c] >= 0) {
                    for (int k = 0; k < tree.nchild[c]; k++) {
                        const int32_t d = tree.child[c] + k;
                        const bool single = tree.child[d] < 0 && tree.count[d] == 1;
                        bm = std::max(bm, single ? dist(&tree.pos[3 * tree.first[d]])
                                                 : dist(&bcen[3 * d]) + bbmax[d]);
                    }
                } else {
                    for (int64_t i = tree.first[c]; i < tree.first[c] + tree.count[c]; i++)
                        bm = std::max(bm, dist(&tree.pos[3 * i]));
                }
                const float half = 0.5f * geom[5 * (size_t)c + 3];
                bbmax[c] = std::min(bm, (double)sqrtf(3.0f * half * half));
            }
        }
        if (tp.mac != 1 || (!geo && !sb))
            return;
        const double bmax0 = 0.5 * tree.size;
        for (int64_t c = 0; c < nn; c++) {
            const double *raw = &braw[35 * (size_t)c];
            const double *cm = &bcen[3 * c];
            const double B2 = raw[raw_index(2, 0, 0)] + raw[raw_index(0, 2, 0)] + raw[raw_index(0, 0, 2)];
            const double B4 = raw[raw_index(4, 0, 0)] + raw[raw_index(0, 4, 0)] +
                              raw[raw_index(0, 0, 4)] + 2 * raw[raw_index(2, 2, 0)] +
                              2 * raw[raw_index(2, 0, 2)] + 2 * raw[raw_index(0, 2, 2)];
            double B6 = 0.0, m5[21] = {0};
            for (int64_t i = tree.first[c]; i < tree.first[c] + tree.count[c]; i++) {
                const double dx = tree.pos[3 * i] - cm[0], dy = tree.pos[3 * i + 1] - cm[1],
                             dz = tree.pos[3 * i + 2] - cm[2];
                const double r2 = dx * dx + dy * dy + dz * dz;
                B6 += tree.mass[i] * r2 * r2 * r2;
                int k = 0;
                for (int a = 5; a >= 0; a--)
                    for (int b = 5 - a; b >= 0; b--, k++)
                        m5[k] += tree.mass[i] * pow(dx, a) * pow(dy, b) * pow(dz, 5 - a - b);
            }
            double *out = &brcrit[3 * (size_t)c];
            if (sb && tree.count[c] > 8) {
                double s3 = 0.0, s5 = 0.0;
                for (int a = 3; a >= 0; a--)
                    for (int b = 3 - a; b >= 0; b--)
                        s3 += multinomial(3, a, b, 3 - a - b) * fabs(raw[raw_index(a, b, 3 - a - b)]);
                int k = 0;
                for (int a = 5; a >= 0; a--)
                    for (int b = 5 - a; b >= 0; b--, k++) s5 += multinomial(5, a, b, 5 - a - b) * fabs(m5[k]);
                const double sb35[2] = {s3 / 9.0, s5 / 27.0};
                arel_rcrit(raw[0], bbmax[c], B2, B4, B6, tree.count[c], tp, bmax0, out, sb35);
            } else {
                arel_rcrit(raw[0], bbmax[c], B2, B4, B6, tree.count[c], tp, bmax0, out);
            }
        }
    }
    static double multinomial(int n, int a, int b, int c) {
        auto f = [](int k) {
            double r = 1;
            for (int i = 2; i <= k; i++) r *= i;
            return r;
        };
        return f(n) / (f(a) * f(b) * f(c));
    }
    // mpole_add_cube (nln/cofm.c:626-648) on raw moments about the centre: a cube
    // of edge a and density rho, its moments taken about its own centre
    static void add_cube(double a, double rho, double *u) {
        const double a3 = a * a * a * rho, a5 = a * a * a3, a7 = a * a * a5;
        u[raw_index(0, 0, 0)] += a3;
        for (int ax = 0; ax < 3; ax++) {
            int e[3] = {0, 0, 0};
            e[ax] = 2;
            u[raw_index(e[0], e[1], e[2])] += a5 / 12.0;
            e[ax] = 4;
            u[raw_index(e[0], e[1], e[2])] += a7 / 80.0;
        }
        u[raw_index(2, 2, 0)] += a7 / 144.0;
        u[raw_index(2, 0, 2)] += a7 / 144.0;
        u[raw_index(0, 2, 2)] += a7 / 144.0;
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
            if (dp.hblock && from.daughters >= dp.min_msink &&
                from.mmterms - from.mmterms_done >= dp.min_msrc)
                mxn_mono(from);
            from.processed = true;
        }
        Sink to = make_sink(to_ent);
        to.mcnt = from.mcnt;
        to.qcnt = from.qcnt;
        to.hcnt = from.hcnt;
        to.qcnt_done = from.qcnt_done;
        to.hcnt_done = from.hcnt_done;
        to.mmcnt = from.mmcnt;
        to.mmcnt_done = from.mmcnt_done;
        to.mmterms = from.mmterms;
        to.mmterms_done = from.mmterms_done;
        to.interactions = from.interactions;
        if (sb) {
            // the near cube (nln/mac.c:745-748, :824-843): fixed at the first cell
            // of <= 32 bodies or of half-size <= 2 mean spacings on the path, or
            // at the body (its leaf cube) if none came before
            const bool ccp = to_ent >= 0 && (cell[to_ent].daughters <= 32 ||
                                             4.0001f * r0 / nx >= ucell[tree.level[to_ent]].halfsz);
            if (ccp && from.clevel == CHUBITS) {
                const float hs = ucell[tree.level[to_ent]].halfsz;
                to.clevel = tree.level[to_ent];
                memcpy(to.cen, cell[to_ent].pos, sizeof to.cen);
                to.cr = 3.0f * hs;
                to.cr2 = sq(3.99f * hs);
            } else {
                to.clevel = from.clevel;
                memcpy(to.cen, from.cen, sizeof to.cen);
                to.cr = from.cr;
                to.cr2 = from.cr2;
            }
            if (to.isbody && from.clevel == CHUBITS) {
                const int64_t leaf = body_leaf[to.first];
                const float *g = &geom[5 * (size_t)leaf];
                for (int a = 0; a < 3; a++) to.cen[a] = g[a] + 0.5f * g[3];
                to.clevel = tree.level[leaf];
                to.cr = 1.5f * g[3];
                to.cr2 = sq(1.995f * g[3]);
            }
            to.near = from.near;
            to.fmass = from.fmass;
            to.done = to.interactions == (int64_t)nimage * tree.n ? true : from.done;
        }
        return to;
    }
    // Per-depth buffers of the walk, reused from sink to sink (the source lists of
    // a periodic walk run to thousands of entries per sink; fresh vectors for
    // every sink spent most of the walk in the allocator and page faults).
    struct Level {
        std::vector<Src> pending, pass, opened;
        std::vector<int> res;
        std::vector<std::pair<int64_t, int64_t>> undo;
    };
    std::vector<Level> lv;  // sized once in run(): a deeper walk never moves a caller's level
    void walk(Sink &s, int32_t ent, const std::vector<Src> &in, size_t depth) {
        if (depth >= lv.size())
            throw std::runtime_error("the sink tree is deeper than the walk's level buffers");
        Level &L = lv[depth];
        std::vector<Src> &pending = L.pending, &pass = L.pass, &opened = L.opened;
        std::vector<int> &res = L.res;
        pending.assign(in.begin(), in.end());
        pass.clear();
        const int64_t h0 = s.hcnt, q0 = s.qcnt, m0 = s.mcnt;
        while (!pending.empty()) {
            if (sb) {
                mac_sb(s, ent, pending, res);
                log_sink_sb(s);
            } else {
                mac(s, ent, pending, res);
            }
            opened.clear();
            for (size_t i = 0; i < pending.size(); i++) {
                if (res[i] == MAC_SPLIT_SINK) {
                    pass.push_back(pending[i]);
                } else if (res[i] == MAC_SPLIT_SRC) {
                    const int32_t e = pending[i].ent;
                    for (int64_t k = child_off[e]; k < child_off[e + 1]; k++)
                        opened.push_back({children[k], pending[i].img, pending[i].bs});
                }
            }
            pending.swap(opened);
        }
        std::vector<std::pair<int64_t, int64_t>> &undo = L.undo;
        undo.clear();
        if (dp.tail_spans) {
            snapshot(0, H, h0, s.hcnt, undo);
            snapshot(1, Q, q0, s.qcnt, undo);
            snapshot(2, M, m0, s.mcnt, undo);
        }
        if (s.isbody) {
            leaf(s, ent_body(ent));
        } else {
            for (int64_t k = child_off[ent]; k < child_off[ent + 1]; k++) {
                Sink d = inherit(s, children[k]);
                walk(d, children[k], pass, depth + 1);
            }
        }
        undo_snapshots(undo);
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
        if (sb || dp.geometric_center)
            setup_geom();
        if (sb)
            setup_sb();
        cell_basis();
        cell.resize(nn);
        rec.assign((size_t)HSZ * nn, 0.0f);
        const bool geo = dp.geometric_center != 0;
        for (int64_t c = 0; c < nn; c++) {
            Cell &C = cell[c];
            const double *raw = &braw[35 * (size_t)c];
            const double *cen = &bcen[3 * c];
            C.mass = (float)raw[0];
            for (int a = 0; a < 3; a++) C.pos[a] = (float)cen[a];
            C.bmax = (float)bbmax[c];
            C.daughters = tree.count[c];
            if (tp.mac == 1) {
                // radii of orders 0, 2, 4 (a cell too small for an order has none: its
                // test never passes -- isquad / ishexa are false for it anyway)
                const double *rc = &brcrit[3 * (size_t)c];
                C.rcrit = (float)rc[0];
                C.rcrit_q = std::isnan(rc[1]) ? (float)rc[0] : (float)rc[1];
                C.rcrit_h = std::isnan(rc[2]) ? C.rcrit_q : (float)rc[2];
            } else {
                // opening angle: the top order P at bmax / theta (the Bmax MAC,
                // nln/cofm.c:568-569) and each lower order k where its truncation
                // bound (b / r)^(k+1) is the top order's, theta^(P+1)
                const int P = tp.order;
                const double b = bbmax[c];
                auto rk = [&](int k) { return (float)(b * pow(tp.theta, -(double)(P + 1) / (k + 1))); };
                C.rcrit = rk(0);
                C.rcrit_q = P >= 2 ? rk(2) : C.rcrit;
                C.rcrit_h = P >= 4 ? rk(4) : C.rcrit_q;
            }
            double r29[HSZ];
            const double R = bbmax[c] > 0 ? bbmax[c] : 1.0;
            raw_to_record<double>(raw, cen, R, r29);
            if (!geo)
                r29[5] = r29[6] = r29[7] = 0.0;  // no dipole about the centre of mass
            if (sb) {
                // the hexacell's unsubtracted umass / uxxxx / uyyyy / uxxyy, then the
                // record of the moments less the cell's cube (nln/cofm.c:345-349, :540-549)
                float *u = &ubs[4 * (size_t)c];
                u[0] = (float)raw[0];
                u[1] = (float)r29[20];
                u[2] = (float)r29[24];
                u[3] = (float)r29[22];
                double sub[35];
                memcpy(sub, raw, sizeof sub);
                add_cube((double)geom[5 * c + 3], -rho0, sub);
                raw_to_record<double>(sub, cen, R, r29);
                if (!geo)
                    r29[5] = r29[6] = r29[7] = 0.0;
            }
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
                    const bool single = tree.child[d] < 0 && tree.count[d] == 1;
                    children.push_back(single ? body_ent(tree.first[d]) : d);
                    if (sb && single)
                        body_leaf[tree.first[d]] = d;
                }
            } else if (tree.count[c] > 1) {
                for (int64_t b = tree.first[c]; b < tree.first[c] + tree.count[c]; b++) {
                    children.push_back(body_ent(b));
                    if (sb)
                        body_leaf[b] = c;
                }
            }
            child_off[c + 1] = (int64_t)children.size();
        }
        tails.assign(6 * (size_t)n, 0);
        inter.assign(n, 0);
        if (dp.tail_spans)
            tail_index.assign(2 * (size_t)n, 0);
        // the root: one source per image (WalkInitSrc / WalkInitSrcPeriodic,
        // nln/mac.c:279-291)
        const int32_t root = tree.child[0] < 0 && tree.count[0] == 1 ? body_ent(0) : 0;
        std::vector<Src> src0;
        for (int i = 0; i < nimage; i++) src0.push_back({root, i, 0});
        Sink s = make_sink(root);
        // one level per tree level (at most 21 below the root) and the bodies
        lv.assign(24, Level());
        walk(s, root, src0, 0);
        qns_flush();
        if (sb)
            mns_flush();
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
    d->subtract_background = 0;
    d->do_pM = 1;
    d->min_msink = 8;
    d->min_msrc = 128;
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
    if (d->subtract_background && p->periodic != 2) {
        fail("pgrav_dl_walk: subtract_background needs the periodic box of explicit images "
             "(periodic 2)");
        return nullptr;
    }
    if (d->min_msink < 0 || d->min_msrc < 0 || !(d->nx >= 0.0f)) {
        fail("pgrav_dl_walk: non-negative M x N limits and nx");
        return nullptr;
    }
    if (d->subtract_background && d->do_pM) {
        // MMvec segments are evaluated with the first body's mass (nln/mac.c:637-639)
        for (int64_t i = 1; i < n; i++)
            if (mass[i] != mass[0]) {
                fail("pgrav_dl_walk: MMvec segments (do_pM) need equal body masses");
                return nullptr;
            }
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
    W->sb = d->subtract_background != 0;
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
    case PGRAV_DL_SOURCES:
        count = (int64_t)W->src.size();
        if (out)
            W->src.copy_out((float *)out);
        return count;
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
    case PGRAV_DL_NODE_GEOM: take(W->geom, 5); break;
    case PGRAV_DL_UBS: take(W->ubs, 4); break;
    case PGRAV_DL_BODY_LEAF: take(W->body_leaf, 1); break;
    case PGRAV_DL_MFLUSH: take(W->mflush, 8); break;
    case PGRAV_DL_MSS_INDEX: take(W->mss_index, 1); break;
    case PGRAV_DL_MSS_SEG: take(W->mss_seg, 2); break;
    case PGRAV_DL_MSS_SOURCE_N: take(W->mss_source_n, 1); break;
    case PGRAV_DL_MSEG: take(W->mseg, 2); break;
    case PGRAV_DL_MCALLS: take(W->mcalls, 6); break;
    case PGRAV_DL_MMTAILS: take(W->mmtails, 3); break;
    case PGRAV_DL_BACKGROUND: take(W->background, 3); break;
    case PGRAV_DL_SB_SINK: take(W->sbsink, 6); break;
    case PGRAV_DL_LOG_SB: take(W->log_sb, 8); break;
    case PGRAV_DL_LOG_FLAGS: take(W->log_flags, 2); break;
    case PGRAV_DL_TAIL_RUNS: take(W->tail_runs, 3); break;
    case PGRAV_DL_TAIL_INDEX: take(W->tail_index, 2); break;
    case PGRAV_DL_SB_STATS:
        p = W->sbstats;
        bytes = sizeof W->sbstats;
        count = 8;
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
