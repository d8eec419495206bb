// Force-evaluation plan: the native runtime around the batched kernels.
//
// pgrav_plan_create() runs the host side (Morton sort, octree, FP64 moments,
// sink groups, interaction lists) and uploads everything once; after that a
// force evaluation is a single kernel launch on device-resident data, which a
// caller may also capture into a hipGraph and replay.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <new>
#include <stdexcept>
#include <vector>

#include "../../include/pgrav.h"
#include "pgrav_batch.h"
#include "pgrav_device.h"
#include "pgrav_ewald.h"
#include "pgrav_host.h"
#include "pgrav_le.h"
#include "pgrav_le_launch.h"
#include "pgrav_moments.h"
#include "pgrav_internal.h"

namespace pgrav {

static const int NQ_PLAIN = 32;  // streams for the one-launch-per-group path (nln/pgrav.cu:811)

struct Plan {
    pgrav_params prm;
    TreeParams tp;
    Tree tree;
    Lists lists;
    std::vector<int64_t> inv;  // original index -> sorted slot
    hipStream_t stream = nullptr;
    hipStream_t qs[NQ_PLAIN] = {};
    hipEvent_t qev[NQ_PLAIN] = {};  // each slot's last launch
    hipEvent_t ev0 = nullptr;
    // the last evaluation, on whichever stream it ran: what writes to the bodies
    // and dispatch tables wait for (evaluations of one plan must not overlap)
    hipEvent_t ev_last = nullptr;
    hipEvent_t ev_join = nullptr;  // orders the plan's stream after a caller's stream
    // device-side body update: the permutation, staging for host arrays, the
    // gathered slots and the first bad body
    int64_t *d_perm = nullptr;
    double *u_pos = nullptr, *u_mass = nullptr;
    double *u_w = nullptr, *u_g = nullptr;  // the host-array adjoint's weights and gradient
    double4 *u_b4 = nullptr;
    unsigned long long *u_bad = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<void *> allocs;
    BatchView v;
    double t_tree = 0, t_lists = 0, t_upload = 0;
    int64_t nnodes = 0;
    // adjoint buffers and transposed lists, set up on the first adjoint
    AdjointView av;
    bool adj_ready = false;
    // FP64 upward pass on the device (pgrav_moments.hip), set up on first use
    MomentsView mv;
    std::vector<int> lev_off;  // nodes of level L: lev_nodes[lev_off[L] .. lev_off[L+1])
    void ensure_moments_view();
    // far-field pass of the explicit-image scheme (periodic = 2)
    double *le_W = nullptr, *le_partial = nullptr, *le_coef = nullptr;
    int le_blocks = 0;

    template <typename P>
    P *dalloc(size_t bytes) {
        void *p = nullptr;
        if (!check(hipMalloc(&p, bytes ? bytes : 16), "hipMalloc"))
            throw std::runtime_error("device allocation failed");
        allocs.push_back(p);
        return (P *)p;
    }
    template <typename P>
    P *upload(const P *h, size_t count) {
        P *d = dalloc<P>(count * sizeof(P));
        if (count && !check(hipMemcpy(d, h, count * sizeof(P), hipMemcpyHostToDevice), "hipMemcpy"))
            throw std::runtime_error("upload failed");
        return d;
    }
    ~Plan() {
        if (gexec)
            (void)hipGraphExecDestroy(gexec);
        if (graph)
            (void)hipGraphDestroy(graph);
        if (stream)
            (void)hipStreamSynchronize(stream);
        for (void *p : allocs) (void)hipFree(p);
        for (auto &s : qs)
            if (s)
                (void)hipStreamDestroy(s);
        if (stream)
            (void)hipStreamDestroy(stream);
        if (ev0)
            (void)hipEventDestroy(ev0);
        for (hipEvent_t e : {ev_last, ev_join})
            if (e)
                (void)hipEventDestroy(e);
        for (auto &e : qev)
            if (e)
                (void)hipEventDestroy(e);
    }
};

static double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Device copies of the tree structure and of the host's moments, so that a
// later update recomputes the moments on the GPU from the moved bodies.
void Plan::ensure_moments_view() {
    if (mv.raw)
        return;
    const Tree &t = tree;
    const int64_t nn = t.nnodes();
    std::vector<int> cnt(t.count.begin(), t.count.end()), ch(t.child.begin(), t.child.end()),
        nch(t.nchild.begin(), t.nchild.end());
    const int NL = 23;
    lev_off.assign(NL + 1, 0);
    for (int64_t i = 0; i < nn; i++) lev_off[t.level[i] + 1]++;
    for (int L = 0; L < NL; L++) lev_off[L + 1] += lev_off[L];
    std::vector<int> lev_nodes(nn), fill(lev_off.begin(), lev_off.end() - 1);
    for (int64_t i = 0; i < nn; i++) lev_nodes[fill[t.level[i]]++] = (int)i;
    std::vector<double> c4((size_t)nn * 4);
    for (int64_t i = 0; i < nn; i++) {
        for (int q = 0; q < 3; q++) c4[4 * i + q] = t.cofm[3 * i + q];
        c4[4 * i + 3] = t.raw[35 * (size_t)i];
    }
    mv.bpos = v.bpos;
    mv.first = upload(t.first.data(), t.first.size());
    mv.count = upload(cnt.data(), cnt.size());
    mv.child = upload(ch.data(), ch.size());
    mv.nchild = upload(nch.data(), nch.size());
    mv.lev_nodes = upload(lev_nodes.data(), lev_nodes.size());
    mv.cpos = (double4 *)v.cpos;
    mv.bmax = upload(t.bmax.data(), t.bmax.size());
    mv.raw = upload(t.raw.data(), t.raw.size());
    mv.mom32 = (float *)v.mom32;
    mv.mom64 = (double *)v.mom64;
}

static int auto_blocks() {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t pr;
        if (hipGetDeviceProperties(&pr, dev) == hipSuccess)
            cus = pr.multiProcessorCount;
    }
    return cus * 32;  // 32 single-wave workgroups per CU: 8 per SIMD, a full CU
}

static hipStream_t stream_of(Plan *P, void *stream) {
    return stream ? (hipStream_t)stream : P->stream;
}

// wait until the last evaluation (on any stream) and the plan's own stream are idle
static bool quiesce(Plan *P) {
    return check(hipEventSynchronize(P->ev_last), "hipEventSynchronize") &&
           check(hipStreamSynchronize(P->stream), "hipStreamSynchronize");
}

// work queued on `before` completes before what is queued on `after` next
static bool order_streams(Plan *P, hipStream_t before, hipStream_t after) {
    if (before == after)
        return true;
    return check(hipEventRecord(P->ev_join, before), "hipEventRecord") &&
           check(hipStreamWaitEvent(after, P->ev_join, 0), "hipStreamWaitEvent");
}

static void ensure_perm(Plan *P) {
    if (!P->d_perm)
        P->d_perm = P->upload(P->tree.perm.data(), P->tree.perm.size());
}

// bodies (caller order, device arrays ready on stream cs) into the sorted slots,
// then the FP64 moments and records recomputed on the GPU; returns when done
static int update_device(Plan *P, const double *pos3, const double *mass, hipStream_t cs) {
    const int64_t n = P->tree.n;
    ensure_perm(P);
    P->ensure_moments_view();
    if (!P->u_b4) {
        P->u_b4 = P->dalloc<double4>(sizeof(double4) * n);
        P->u_bad = P->dalloc<unsigned long long>(sizeof(unsigned long long));
    }
    unsigned long long bad = 0;
    if (!quiesce(P) || !order_streams(P, cs, P->stream) ||
        !check(launch_gather_bodies(pos3, mass, P->d_perm, n, P->prm.periodic != 0, P->prm.box,
                                    P->v.bpos, P->u_b4, P->u_bad, P->stream), "body gather kernel") ||
        !check(hipMemcpyAsync(&bad, P->u_bad, sizeof(bad), hipMemcpyDeviceToHost, P->stream), "hipMemcpyAsync") ||
        !check(hipStreamSynchronize(P->stream), "hipStreamSynchronize"))
        return -1;
    if (bad != ~0ull) {
        fail("pgrav_plan_update: body %llu has a non-positive mass or a non-finite value", bad);
        return -1;
    }
    if (!check(hipMemcpyAsync((void *)P->v.bpos, P->u_b4, sizeof(double4) * n,
                              hipMemcpyDeviceToDevice, P->stream), "hipMemcpyAsync") ||
        !check(launch_moments(P->mv, P->lev_off.data(), (int)P->lev_off.size() - 1, P->nnodes,
                              P->prm.order, P->stream), "moment kernels") ||
        !check(hipStreamSynchronize(P->stream), "hipStreamSynchronize"))
        return -1;
    return 0;
}

// the far images and background of periodic = 2 plans, added to the result
static int far_field(Plan *P, int variant, hipStream_t st) {
    if (P->prm.periodic != 2)
        return 0;
    const bool f64 = variant == VAR_F64 || variant == VAR_PLAIN64;
    return check(launch_ewald_le(P->v.bpos, P->tree.n, P->prm.box, P->prm.G, P->le_W, P->le_partial,
                                 P->le_blocks, P->le_coef, f64 ? nullptr : P->v.out32,
                                 f64 ? P->v.out64 : nullptr, st), "far-field kernels")
               ? 0 : -1;
}

static int evaluate(Plan *P, int variant, hipStream_t st) {
    if (variant == VAR_PLAIN64) {
        if (!P->v.mom64) {
            fail("the FP64 tables were not built (pgrav_params.fp64 = 0)");
            return -1;
        }
        // one launch per group over 32 slot streams, as the reference's slot queue
        // (nln/pgrav.cu:811-851, qallocCUDA): a slot takes its next group only once its
        // previous launch has completed, so at most NQ_PLAIN launches are in flight, as
        // there; the host waits on the oldest slot where the reference spins over all 32
        if (!check(hipEventRecord(P->ev0, st), "hipEventRecord"))
            return -1;
        for (int q = 0; q < NQ_PLAIN; q++)
            if (!check(hipStreamWaitEvent(P->qs[q], P->ev0, 0), "hipStreamWaitEvent"))
                return -1;
        for (int g = 0; g < P->v.ngroups; g++) {
            const int q = g % NQ_PLAIN;
            if (g >= NQ_PLAIN && !check(hipEventSynchronize(P->qev[q]), "hipEventSynchronize"))
                return -1;
            if (!check(launch_group64(P->v, g, P->qs[q]), "group kernel") ||
                !check(hipEventRecord(P->qev[q], P->qs[q]), "hipEventRecord"))
                return -1;
        }
        for (int q = 0; q < NQ_PLAIN && q < P->v.ngroups; q++)
            if (!check(hipStreamWaitEvent(st, P->qev[q], 0), "hipStreamWaitEvent"))
                return -1;
        return far_field(P, variant, st);
    }
    if (variant == VAR_F64 && !P->v.mom64) {
        fail("the FP64 tables were not built (pgrav_params.fp64 = 0)");
        return -1;
    }
    if (variant < 0 || variant > VAR_F64) {
        fail("unknown variant %d", variant);
        return -1;
    }
    int pb = P->prm.persistent_blocks;
    if (pb < 0)
        pb = auto_blocks();
    if (!check(launch_batch(P->v, variant, pb, st), "batched kernel"))
        return -1;
    return far_field(P, variant, st);
}

}  // namespace pgrav

using namespace pgrav;

extern "C" {

void pgrav_params_default(pgrav_params *p) {
    memset(p, 0, sizeof(*p));
    p->theta = 0.6;
    p->group_max = 64;
    p->leaf_max = 16;
    p->order = 4;
    p->pcut = 0;
    p->soft_type = 1;
    p->eps = 1e-4;
    p->G = 1.0;
    p->periodic = 0;
    p->box = 1.0;
    p->nthreads = 0;
    p->merge_groups = 1;
    p->persistent_blocks = -1;  // persistent grid, 32 workgroups per CU
    p->fp64 = 1;
    p->dispatch = 2;
    p->mixed_order = 0;
    p->nimage = 2;  // nln/main.c:246
    p->rcb_cell = 32768;
    p->mac = 0;
    p->quad_ncut = 7;   // nln/main.c:540
    p->hexa_ncut = 20;  // nln/main.c:541
    p->ewald_ratio = 0.05;
}

void *pgrav_plan_create(const double *pos3, const double *mass, int64_t n, const pgrav_params *p) {
    clear_error();
    if (!pos3 || !mass || n <= 0 || !p) {
        fail("pgrav_plan_create: bad arguments");
        return nullptr;
    }
    if (p->soft_type != SOFT_NONE && p->soft_type != SOFT_PLUMMER && p->soft_type != SOFT_K1) {
        fail("soft_type must be 0 (none), 1 (Plummer) or 4 (K1), nln/mac.c:196-237");
        return nullptr;
    }
    if (p->body_unroll != 0 && p->body_unroll != 2 && p->body_unroll != 4 && p->body_unroll != 8) {
        fail("pgrav_plan_create: body_unroll must be 0, 2, 4 or 8 (got %d)", p->body_unroll);
        return nullptr;
    }
    if (!(p->ewald_ratio >= 0)) {
        fail("pgrav_plan_create: ewald_ratio must be >= 0 (got %g)", p->ewald_ratio);
        return nullptr;
    }
    if (p->soft_type != SOFT_NONE && !(p->eps > 0)) {
        fail("pgrav_plan_create: softening needs eps > 0 (got %g)", p->eps);
        return nullptr;
    }
    TreeParams tp;
    const std::string bad = tree_params_from(p, tp);
    if (!bad.empty()) {
        fail("pgrav_plan_create: %s", bad.c_str());
        return nullptr;
    }
    if (n > (int64_t)UINT32_MAX) {
        fail("more than 2^32 bodies");
        return nullptr;
    }
    int ndev = 0;
    if (!check(hipGetDeviceCount(&ndev), "hipGetDeviceCount"))
        return nullptr;
    Plan *P = new (std::nothrow) Plan();
    if (!P) {
        fail("out of host memory");
        return nullptr;
    }
    try {
        P->prm = *p;
        P->tp = tp;
        double t0 = now();
        build_tree(pos3, mass, n, P->tp, P->tree);
        double t1 = now();
        build_groups(P->tree, P->tp, P->lists);
        build_lists(P->tree, P->tp, P->lists);
        double t2 = now();
        P->t_tree = t1 - t0;
        P->t_lists = t2 - t1;
        P->nnodes = P->tree.nnodes();
        P->inv.resize(n);
        for (int64_t i = 0; i < n; i++) P->inv[P->tree.perm[i]] = i;

        if (!check(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking), "hipStreamCreate") ||
            !check(hipEventCreate(&P->ev0), "hipEventCreate") ||
            !check(hipEventCreateWithFlags(&P->ev_last, hipEventDisableTiming), "hipEventCreate") ||
            !check(hipEventCreateWithFlags(&P->ev_join, hipEventDisableTiming), "hipEventCreate") ||
            !check(hipEventRecord(P->ev_last, P->stream), "hipEventRecord"))
            throw std::runtime_error("stream");
        for (int q = 0; q < NQ_PLAIN; q++)
            if (!check(hipStreamCreateWithFlags(&P->qs[q], hipStreamNonBlocking), "hipStreamCreate") ||
                !check(hipEventCreateWithFlags(&P->qev[q], hipEventDisableTiming), "hipEventCreate"))
                throw std::runtime_error("stream");
        // bodies
        std::vector<double> b4(4 * (size_t)n);
        for (int64_t i = 0; i < n; i++) {
            b4[4 * i + 0] = P->tree.pos[3 * i + 0];
            b4[4 * i + 1] = P->tree.pos[3 * i + 1];
            b4[4 * i + 2] = P->tree.pos[3 * i + 2];
            b4[4 * i + 3] = P->tree.mass[i];
        }
        P->v.bpos = (const double4 *)P->upload(b4.data(), b4.size());
        std::vector<double>().swap(b4);
        // cells
        std::vector<float> r32;
        std::vector<double> r64, c4;
        prepare_cells(P->tree, p->order, r32, p->fp64 ? &r64 : nullptr, c4);
        P->v.cpos = (const double4 *)P->upload(c4.data(), c4.size());
        P->v.mom32 = P->upload(r32.data(), r32.size());
        if (p->fp64)
            P->v.mom64 = P->upload(r64.data(), r64.size());
        // groups and lists
        Lists &L = P->lists;
        P->v.ngroups = (int)L.ngroups;
        P->v.gfirst = P->upload(L.gfirst.data(), L.gfirst.size());
        P->v.gcount = P->upload(L.gcount.data(), L.gcount.size());
        if (!L.gslots.empty()) {
            std::vector<uint32_t> slots(L.gslots.begin(), L.gslots.end());
            P->v.gslot_off = P->upload(L.gslot_off.data(), L.gslot_off.size());
            P->v.gslot = P->upload(slots.data(), slots.size());
        }
        P->v.gcen = (const double4 *)P->upload(L.gcenter.data(), L.gcenter.size());
        P->v.lpt = P->upload(L.lpt.data(), L.lpt.size());
        P->v.xoff = P->upload(L.xoff.data(), L.xoff.size());
        P->v.order_np = P->upload(L.order_np.data(), L.order_np.size());
        {
            // one CSR over all kinds: offsets [NLIST][G+1] into one index array
            const size_t G1 = (size_t)L.ngroups + 1;
            std::vector<int64_t> off(NLIST * G1);
            size_t total = 0;
            for (int k = 0; k < NLIST; k++) {
                for (size_t g = 0; g < G1; g++) off[k * G1 + g] = L.off[k][g] + (int64_t)total;
                total += L.idx[k].size();
            }
            P->v.off = P->upload(off.data(), off.size());
            uint32_t *idx = P->dalloc<uint32_t>(std::max<size_t>(1, total) * sizeof(uint32_t));
            size_t at = 0;
            for (int k = 0; k < NLIST; k++) {
                const size_t m = L.idx[k].size();
                if (m && !check(hipMemcpy(idx + at, L.idx[k].data(), m * sizeof(uint32_t),
                                          hipMemcpyHostToDevice), "hipMemcpy"))
                    throw std::runtime_error("upload failed");
                at += m;
                std::vector<uint32_t>().swap(L.idx[k]);  // the device copy is the one used
            }
            P->v.idx = idx;
            if (p->periodic == 2) {
                uint8_t *img = P->dalloc<uint8_t>(std::max<size_t>(1, total));
                at = 0;
                for (int k = 0; k < NLIST; k++) {
                    const size_t m = L.img[k].size();
                    if (m && !check(hipMemcpy(img + at, L.img[k].data(), m, hipMemcpyHostToDevice),
                                    "hipMemcpy"))
                        throw std::runtime_error("upload failed");
                    at += m;
                    std::vector<uint8_t>().swap(L.img[k]);
                }
                P->v.img = img;
            }
        }
        P->v.out32 = P->dalloc<float4>(sizeof(float4) * (size_t)n);
        if (p->fp64)
            P->v.out64 = P->dalloc<double4>(sizeof(double4) * (size_t)n);
        P->v.counter = P->dalloc<int>(64);
        if (!check(hipMemset(P->v.counter, 0, 64), "hipMemset"))
            throw std::runtime_error("memset");
        P->v.soft = p->soft_type;
        P->v.unroll = p->body_unroll ? p->body_unroll : 4;
        P->v.ewratio = p->ewald_ratio;
        P->v.eps = p->eps;
        P->v.G = p->G;
        P->v.periodic = p->periodic;
        P->v.nimage = p->nimage;
        P->v.box = p->box;
        if (p->periodic == 2) {
            std::vector<double> W, Wt((size_t)LE_NMON * LE_NMON);
            le_matrix(p->nimage, W);
            for (int b = 0; b < LE_NMON; b++)
                for (int g = 0; g < LE_NMON; g++) Wt[(size_t)g * LE_NMON + b] = W[(size_t)b * LE_NMON + g];
            P->le_W = P->upload(Wt.data(), Wt.size());
            P->le_blocks = (int)std::min<int64_t>(2048, (n + 511) / 512);
            P->le_partial = P->dalloc<double>(sizeof(double) * LE_NMON * (P->le_blocks + 1));
            P->le_coef = P->dalloc<double>(sizeof(double) * LE_NMON);
        }
        if (p->periodic) {
            const size_t ne = (size_t)EW_REC * EWALD_NODES;  // pgrav_ewald.h
            double *t64 = P->dalloc<double>(ne * sizeof(double));
            float *t32 = P->dalloc<float>(ne * sizeof(float));
            if (!check(launch_ewald_table(p->box, t64, t32, P->stream), "Ewald table kernel"))
                throw std::runtime_error("ewald");
            P->v.ewald32 = t32;
            P->v.ewald64 = t64;
        }
        if (!check(hipStreamSynchronize(P->stream), "hipStreamSynchronize"))
            throw std::runtime_error("sync");
        P->t_upload = now() - t2;
    } catch (const std::exception &e) {
        if (!failed())
            fail("pgrav_plan_create: %s", e.what());
        delete P;
        return nullptr;
    }
    return P;
}

void pgrav_plan_destroy(void *plan) { delete (Plan *)plan; }

int pgrav_plan_update(void *plan, const double *pos3, const double *mass) {
    Plan *P = (Plan *)plan;
    clear_error();
    if (!P || !pos3 || !mass) {
        fail("pgrav_plan_update: null argument");
        return -1;
    }
    try {
        // the caller's arrays go to the device as they are; the bodies move into
        // their sorted slots there and the moments, centres, bmax and records are
        // recomputed from them in FP64
        const int64_t n = P->tree.n;
        if (!P->u_pos) {
            P->u_pos = P->dalloc<double>(sizeof(double) * 3 * n);
            P->u_mass = P->dalloc<double>(sizeof(double) * n);
        }
        if (!quiesce(P) ||
            !check(hipMemcpyAsync(P->u_pos, pos3, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
                                  P->stream), "hipMemcpyAsync") ||
            !check(hipMemcpyAsync(P->u_mass, mass, sizeof(double) * n, hipMemcpyHostToDevice,
                                  P->stream), "hipMemcpyAsync"))
            return -1;
        return update_device(P, P->u_pos, P->u_mass, P->stream);
    } catch (const std::exception &e) {
        fail("pgrav_plan_update: %s", e.what());
        return -1;
    }
}

int pgrav_plan_update_device(void *plan, const double *pos3, const double *mass, void *stream) {
    Plan *P = (Plan *)plan;
    clear_error();
    if (!P || !pos3 || !mass) {
        fail("pgrav_plan_update_device: null argument");
        return -1;
    }
    try {
        return update_device(P, pos3, mass, stream_of(P, stream));
    } catch (const std::exception &e) {
        fail("pgrav_plan_update_device: %s", e.what());
        return -1;
    }
}

// The adjoint's buffers and the transposed lists, built once per plan from the
// device's lists (the lists are fixed for the plan's lifetime).
static int setup_adjoint(Plan *P) {
    const int64_t n = P->tree.n, nn = P->nnodes;
    const Tree &t = P->tree;
    const Lists &L = P->lists;
    const int64_t G = L.ngroups;
    AdjointView &A = P->av;
    int64_t ncell = 0;
    for (int k = 0; k < L_BODY; k++) ncell += L.off[k][G];
    const int64_t nbody = L.off[L_BODY][G];
    if (ncell + nbody >= ((int64_t)1 << 32)) {
        fail("pgrav_plan_adjoint: more than 2^32 list entries");
        return -1;
    }
    std::vector<uint32_t> idx(std::max<int64_t>(1, ncell + nbody));
    if (ncell + nbody &&
        (!check(hipMemcpyAsync(idx.data(), P->v.idx, sizeof(uint32_t) * (ncell + nbody),
                               hipMemcpyDeviceToHost, P->stream), "hipMemcpyAsync") ||
         !check(hipStreamSynchronize(P->stream), "hipStreamSynchronize")))
        return -1;
    // raw moments -> prepared record is linear and centre-independent (R = 1)
    std::vector<float> mat(46 * 35);
    for (int a = 0; a < 35; a++) {
        double raw[35] = {0}, c[3] = {0, 0, 0}, q[29], pr[PREC4];
        raw[a] = 1.0;
        raw_to_record<double>(raw, c, 1.0, q);
        q[5] = q[6] = q[7] = 0.0;
        prep_record<double>(q, 4, 0, pr);
        for (int k = 0; k < 46; k++) mat[k * 35 + a] = (float)pr[k];
    }
    std::vector<int> leaf(n, -1), parent(t.parent.begin(), t.parent.end());
    for (int64_t c = 0; c < nn; c++)
        if (t.child[c] < 0)
            for (int64_t i = t.first[c]; i < t.first[c] + t.count[c]; i++) leaf[i] = (int)c;
    // transposed cell lists: for each cell the groups listing it, in group
    // (Morton) order, so that the cells processed together walk neighbouring
    // groups and every sum has a fixed order; work items of <= 128 groups
    std::vector<int64_t> toff(nn + 1, 0);
    for (int64_t e = 0; e < ncell; e++) toff[idx[e] + 1]++;
    for (int64_t c = 0; c < nn; c++) toff[c + 1] += toff[c];
    std::vector<int> tcells;
    std::vector<longlong2> trange;
    std::vector<int64_t> cchunk(nn + 1, 0);
    for (int64_t c = 0; c < nn; c++) {
        for (int64_t a = toff[c]; a < toff[c + 1]; a += 128) {
            tcells.push_back((int)c);
            trange.push_back(make_longlong2(a, std::min<int64_t>(toff[c + 1], a + 128)));
        }
        cchunk[c + 1] = (int64_t)tcells.size();
    }
    std::vector<int64_t> cursor(toff.begin(), toff.end() - 1);
    std::vector<uint4> tl(std::max<int64_t>(1, ncell));
    int64_t base = 0;
    for (int k = 0; k < L_BODY; k++) {
        for (int64_t g = 0; g < G; g++) {
            // a run's first sink, or a body-set group's index (its slot table)
            const uint32_t ref = (uint32_t)(P->v.gslot ? g : L.gfirst[g]);
            const uint32_t ck = (uint32_t)L.gcount[g] | ((uint32_t)k << 8);
            for (int64_t e = base + L.off[k][g]; e < base + L.off[k][g + 1]; e++)
                tl[cursor[idx[e]]++] = make_uint4(ref, ck, (uint32_t)e, (uint32_t)((uint64_t)e >> 32));
        }
        base += L.off[k][G];
    }
    // group order within a cell across kinds (a cell has one kind per group)
    for (int64_t c = 0; c < nn; c++)
        if (toff[c + 1] - toff[c] > 1)
            std::sort(tl.begin() + toff[c], tl.begin() + toff[c + 1],
                      [](const uint4 &a, const uint4 &b) { return a.x != b.x ? a.x < b.x : a.z < b.z; });
    // transposed body lists, per leaf: a body list names runs of consecutive
    // sorted bodies, so a run's part inside one leaf is a segment (first entry
    // position, first and end body offset in the leaf), and the leaf's bodies
    // read their entries of that segment from consecutive positions; a leaf's
    // segments in ascending list position
    std::vector<int64_t> lfirst(nn, 0), loff(nn + 1, 0);
    for (int64_t c = 0; c < nn; c++) lfirst[c] = t.first[c];
    std::vector<uint4> seg;
    std::vector<int> seg_leaf;
    for (int64_t e = 0; e < nbody; e++) {
        const int64_t j = idx[ncell + e];
        const int c = leaf[j];
        const uint32_t u = (uint32_t)(j - t.first[c]);
        if (e && seg_leaf.back() == c && (int64_t)idx[ncell + e - 1] == j - 1 && seg.back().z == u) {
            seg.back().z = u + 1;
            continue;
        }
        seg.push_back(make_uint4((uint32_t)e, u, u + 1, 0u));
        seg_leaf.push_back(c);
    }
    for (int c : seg_leaf) loff[c + 1]++;
    for (int64_t c = 0; c < nn; c++) loff[c + 1] += loff[c];
    std::vector<uint4> lseg(std::max<size_t>(1, seg.size()));
    {
        std::vector<int64_t> lcur(loff.begin(), loff.end() - 1);
        for (size_t s = 0; s < seg.size(); s++) lseg[lcur[seg_leaf[s]]++] = seg[s];
    }
    // a body's segments in chunks of max(16, ceil(segments / 64)) (the device
    // computes the same length from loff)
    std::vector<int64_t> pstart(n + 1, 0);
    for (int64_t j = 0; j < n; j++) {
        const int64_t ns = loff[leaf[j] + 1] - loff[leaf[j]];
        const int64_t len = std::max<int64_t>(16, (ns + 63) / 64);
        pstart[j + 1] = pstart[j] + (ns + len - 1) / len;
    }
    std::vector<int> pbody(std::max<int64_t>(1, pstart[n]));
    for (int64_t j = 0; j < n; j++)
        for (int64_t t = pstart[j]; t < pstart[j + 1]; t++) pbody[t] = (int)j;

    A.wv = P->dalloc<float4>(sizeof(float4) * n);
    A.gsink = P->dalloc<float4>(sizeof(float4) * n);
    A.gbody = P->dalloc<float4>(sizeof(float4) * n);
    A.gpair = P->dalloc<float4>(sizeof(float4) * std::max<int64_t>(1, ncell));
    A.gentry = P->dalloc<float4>(sizeof(float4) * std::max<int64_t>(1, nbody));
    A.ncell_entries = ncell;
    A.tcells = P->upload(tcells.data(), tcells.size());
    A.ntcells = (int)tcells.size();
    A.trange = P->upload(trange.data(), trange.size());
    A.tlist = P->upload(tl.data(), tl.size());
    A.gchunk = P->dalloc<float>(sizeof(float) * 52 * std::max<size_t>(1, tcells.size()));
    A.cchunk_off = P->upload(cchunk.data(), cchunk.size());
    A.loff = P->upload(loff.data(), loff.size());
    A.lseg = P->upload(lseg.data(), lseg.size());
    A.lfirst = P->upload(lfirst.data(), lfirst.size());
    A.pstart = P->upload(pstart.data(), pstart.size());
    A.pbody = P->upload(pbody.data(), pbody.size());
    A.nbchunks = pstart[n];
    A.gbchunk = P->dalloc<float4>(sizeof(float4) * std::max<int64_t>(1, pstart[n]));
    A.raw = P->mv.raw;
    A.mat = P->upload(mat.data(), mat.size());
    A.rix = raw_index();
    A.nnodes = nn;
    A.gnode = P->dalloc<float>(sizeof(float) * 40 * (size_t)nn);
    A.leaf_of = P->upload(leaf.data(), leaf.size());
    A.parent = P->upload(parent.data(), parent.size());
    A.n = n;
    A.out = P->dalloc<double4>(sizeof(double4) * n);
    P->adj_ready = true;
    return 0;
}

// the adjoint on device arrays in caller order: weights w4 (float or double
// quadruples) ready on stream cs, gradient grad4 (double quadruples) queued on cs
static bool aligned_quads(const void *p, int fp64, const char *what) {
    if ((uintptr_t)p % (fp64 ? 32 : 16)) {
        fail("%s: the array must be aligned to one quadruple (%d bytes)", what, fp64 ? 32 : 16);
        return false;
    }
    return true;
}

static int adjoint_device(Plan *P, const void *w4, int w_fp64, double *grad4, hipStream_t cs) {
    if (!aligned_quads(w4, w_fp64, "adjoint weights") || !aligned_quads(grad4, 1, "adjoint gradient"))
        return -1;
    if (P->prm.periodic) {
        fail("the adjoint covers the isolated kernels only");
        return -1;
    }
    P->ensure_moments_view();  // the device raw moments are the current ones
    ensure_perm(P);
    if (!P->adj_ready && setup_adjoint(P))
        return -1;
    AdjointView &A = P->av;
    const int64_t n = P->tree.n;
    if (!order_streams(P, cs, P->stream) ||
        !check(launch_gather4(w4, w_fp64, P->d_perm, n, (float4 *)A.wv, P->stream), "weight gather") ||
        !check(launch_adjoint(P->v, A, P->stream), "adjoint kernels") ||
        !check(launch_scatter4(A.out, 1, P->d_perm, n, grad4, 1, P->stream), "gradient scatter") ||
        !order_streams(P, P->stream, cs))
        return -1;
    return 0;
}

int pgrav_plan_adjoint(void *plan, const double *w4, double *grad4) {
    Plan *P = (Plan *)plan;
    clear_error();
    if (!P || !w4 || !grad4) {
        fail("pgrav_plan_adjoint: null argument");
        return -1;
    }
    try {
        const int64_t n = P->tree.n;
        if (!P->u_w) {
            P->u_w = P->dalloc<double>(sizeof(double) * 4 * n);
            P->u_g = P->dalloc<double>(sizeof(double) * 4 * n);
        }
        if (!check(hipMemcpyAsync(P->u_w, w4, sizeof(double) * 4 * n, hipMemcpyHostToDevice,
                                  P->stream), "hipMemcpyAsync") ||
            adjoint_device(P, P->u_w, 1, P->u_g, P->stream) ||
            !check(hipMemcpyAsync(grad4, P->u_g, sizeof(double) * 4 * n, hipMemcpyDeviceToHost,
                                  P->stream), "hipMemcpyAsync") ||
            !check(hipStreamSynchronize(P->stream), "hipStreamSynchronize"))
            return -1;
    } catch (const std::exception &e) {
        fail("pgrav_plan_adjoint: %s", e.what());
        return -1;
    }
    return 0;
}

int pgrav_plan_adjoint_device(void *plan, const void *w4, int w_fp64, double *grad4, void *stream) {
    Plan *P = (Plan *)plan;
    clear_error();
    if (!P || !w4 || !grad4) {
        fail("pgrav_plan_adjoint_device: null argument");
        return -1;
    }
    try {
        return adjoint_device(P, w4, w_fp64 ? 1 : 0, grad4, stream_of(P, stream));
    } catch (const std::exception &e) {
        fail("pgrav_plan_adjoint_device: %s", e.what());
        return -1;
    }
}

int pgrav_plan_stats(void *plan, pgrav_stats *st) {
    Plan *P = (Plan *)plan;
    if (!P || !st) {
        fail("pgrav_plan_stats: null argument");
        return -1;
    }
    memset(st, 0, sizeof(*st));
    st->n = P->tree.n;
    st->nnodes = P->nnodes;
    st->ngroups = P->lists.ngroups;
    for (int k = 0; k < NLIST; k++) {
        st->pairs[k] = P->lists.pairs[k];
        st->entries[k] = P->lists.off[k].back();
    }
    st->t_tree = P->t_tree;
    st->t_lists = P->t_lists;
    st->t_upload = P->t_upload;
    st->mean_group = P->lists.ngroups ? (double)P->tree.n / P->lists.ngroups : 0.0;
    return 0;
}

int pgrav_plan_set_dispatch(void *plan, int mode) {
    Plan *P = (Plan *)plan;
    clear_error();
    if (!P || mode < 0 || mode > 3) {
        fail("pgrav_plan_set_dispatch: bad arguments");
        return -1;
    }
    Lists &L = P->lists;
    const int64_t G = L.ngroups;
    std::vector<double> cost(G);
    for (int64_t g = 0; g < G; g++)
        cost[g] = wave_cost(L.off[L_HEXA][g + 1] - L.off[L_HEXA][g], L.off[L_QUAD][g + 1] - L.off[L_QUAD][g],
                            L.off[L_MONO][g + 1] - L.off[L_MONO][g], L.off[L_BODY][g + 1] - L.off[L_BODY][g]);
    make_dispatch(cost, mode, L);
    P->prm.dispatch = mode;
    return (quiesce(P) &&
            check(hipMemcpy((void *)P->v.lpt, L.lpt.data(), 4 * G, hipMemcpyHostToDevice), "hipMemcpy") &&
            check(hipMemcpy((void *)P->v.xoff, L.xoff.data(), 4 * L.xoff.size(), hipMemcpyHostToDevice), "hipMemcpy") &&
            check(hipMemcpy((void *)P->v.order_np, L.order_np.data(), 4 * G, hipMemcpyHostToDevice), "hipMemcpy"))
               ? 0 : -1;
}

int pgrav_plan_set_launch(void *plan, int persistent_blocks) {
    Plan *P = (Plan *)plan;
    if (!P) {
        fail("null plan");
        return -1;
    }
    P->prm.persistent_blocks = persistent_blocks;
    return 0;
}

int pgrav_plan_set_unroll(void *plan, int unroll) {
    Plan *P = (Plan *)plan;
    clear_error();
    if (!P || (unroll != 2 && unroll != 4 && unroll != 8)) {
        fail("pgrav_plan_set_unroll: the unroll factor must be 2, 4 or 8");
        return -1;
    }
    P->v.unroll = unroll;
    return 0;
}

int pgrav_plan_set_tile(void *plan, int tile) {
    Plan *P = (Plan *)plan;
    clear_error();
    if (!P || (tile != 32 && tile != 64 && tile != 128)) {
        fail("pgrav_plan_set_tile: the tile must be 32, 64 or 128 list entries");
        return -1;
    }
    P->v.tile = tile;
    return 0;
}

int pgrav_plan_evaluate(void *plan, int variant, void *stream) {
    Plan *P = (Plan *)plan;
    clear_error();
    if (!P) {
        fail("null plan");
        return -1;
    }
    hipStream_t st = stream_of(P, stream);
    if (evaluate(P, variant, st))
        return -1;
    return check(hipEventRecord(P->ev_last, st), "hipEventRecord") ? 0 : -1;
}

int pgrav_plan_time(void *plan, int variant, int reps, void *stream, double *ms_per_launch) {
    Plan *P = (Plan *)plan;
    clear_error();
    if (!P || reps < 1 || !ms_per_launch) {
        fail("pgrav_plan_time: bad arguments");
        return -1;
    }
    hipStream_t st = stream_of(P, stream);
    hipEvent_t a, b;
    if (!check(hipEventCreate(&a), "hipEventCreate") || !check(hipEventCreate(&b), "hipEventCreate"))
        return -1;
    int rc = check(hipEventRecord(a, st), "hipEventRecord") ? 0 : -1;
    for (int r = 0; r < reps && rc == 0; r++) rc = evaluate(P, variant, st);
    if (rc == 0 && !check(hipEventRecord(P->ev_last, st), "hipEventRecord"))
        rc = -1;
    if (rc == 0 && check(hipEventRecord(b, st), "hipEventRecord") &&
        check(hipEventSynchronize(b), "hipEventSynchronize")) {
        float ms = 0;
        if (check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime"))
            *ms_per_launch = (double)ms / reps;
        else
            rc = -1;
    } else {
        rc = -1;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return rc;
}

int pgrav_plan_graph_capture(void *plan, int variant) {
    Plan *P = (Plan *)plan;
    clear_error();
    if (!P) {
        fail("null plan");
        return -1;
    }
    if (variant == VAR_PLAIN64) {
        fail("the one-launch-per-group path is not captured");
        return -1;
    }
    if (P->gexec) {
        (void)hipGraphExecDestroy(P->gexec);
        P->gexec = nullptr;
    }
    if (P->graph) {
        (void)hipGraphDestroy(P->graph);
        P->graph = nullptr;
    }
    if (!quiesce(P) ||
        !check(hipStreamBeginCapture(P->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture"))
        return -1;
    int rc = evaluate(P, variant, P->stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(P->stream, &g);
    if (rc || !check(e, "hipStreamEndCapture"))
        return -1;
    P->graph = g;
    return check(hipGraphInstantiate(&P->gexec, P->graph, nullptr, nullptr, 0), "hipGraphInstantiate") ? 0 : -1;
}

int pgrav_plan_graph_launch(void *plan, void *stream) {
    Plan *P = (Plan *)plan;
    clear_error();
    if (!P || !P->gexec) {
        fail("no captured graph");
        return -1;
    }
    hipStream_t st = stream_of(P, stream);
    return check(hipGraphLaunch(P->gexec, st), "hipGraphLaunch") &&
                   check(hipEventRecord(P->ev_last, st), "hipEventRecord")
               ? 0 : -1;
}

int pgrav_plan_synchronize(void *plan) {
    Plan *P = (Plan *)plan;
    if (!P) {
        fail("null plan");
        return -1;
    }
    return quiesce(P) ? 0 : -1;
}

int pgrav_plan_result(void *plan, int variant, double *acc4) {
    Plan *P = (Plan *)plan;
    clear_error();
    if (!P || !acc4) {
        fail("pgrav_plan_result: null argument");
        return -1;
    }
    const int64_t n = P->tree.n;
    if (!check(hipStreamSynchronize(P->stream), "hipStreamSynchronize") ||
        !check(hipDeviceSynchronize(), "hipDeviceSynchronize"))
        return -1;
    if (variant == VAR_F64 || variant == VAR_PLAIN64) {
        if (!P->v.out64) {
            fail("no FP64 result buffer");
            return -1;
        }
        std::vector<double> h(4 * (size_t)n);
        if (!check(hipMemcpy(h.data(), P->v.out64, h.size() * 8, hipMemcpyDeviceToHost), "hipMemcpy"))
            return -1;
        for (int64_t i = 0; i < n; i++)
            for (int k = 0; k < 4; k++) acc4[4 * P->tree.perm[i] + k] = h[4 * i + k];
    } else {
        std::vector<float> h(4 * (size_t)n);
        if (!check(hipMemcpy(h.data(), P->v.out32, h.size() * 4, hipMemcpyDeviceToHost), "hipMemcpy"))
            return -1;
        for (int64_t i = 0; i < n; i++)
            for (int k = 0; k < 4; k++) acc4[4 * P->tree.perm[i] + k] = h[4 * i + k];
    }
    return 0;
}

int pgrav_plan_result_device(void *plan, int variant, void *acc4, int fp64, void *stream) {
    Plan *P = (Plan *)plan;
    clear_error();
    if (!P || !acc4) {
        fail("pgrav_plan_result_device: null argument");
        return -1;
    }
    if (!aligned_quads(acc4, fp64, "pgrav_plan_result_device"))
        return -1;
    const bool f64 = variant == VAR_F64 || variant == VAR_PLAIN64;
    const void *src = f64 ? (const void *)P->v.out64 : (const void *)P->v.out32;
    if (!src) {
        fail("no FP64 result buffer");
        return -1;
    }
    try {
        ensure_perm(P);
        hipStream_t cs = stream_of(P, stream);
        // after the last evaluation, on the caller's stream
        return check(hipStreamWaitEvent(cs, P->ev_last, 0), "hipStreamWaitEvent") &&
                       check(launch_scatter4(src, f64, P->d_perm, P->tree.n, acc4, fp64 ? 1 : 0, cs),
                             "result scatter")
                   ? 0 : -1;
    } catch (const std::exception &e) {
        fail("pgrav_plan_result_device: %s", e.what());
        return -1;
    }
}

int pgrav_plan_direct(void *plan, const int64_t *sample, int ns, double *out4) {
    Plan *P = (Plan *)plan;
    clear_error();
    if (!P || !sample || ns <= 0 || !out4) {
        fail("pgrav_plan_direct: bad arguments");
        return -1;
    }
    std::vector<int64_t> s(ns);
    for (int i = 0; i < ns; i++) {
        if (sample[i] < 0 || sample[i] >= P->tree.n) {
            fail("sample index %lld out of range", (long long)sample[i]);
            return -1;
        }
        s[i] = P->inv[sample[i]];
    }
    const int blocks = (ns + 63) / 64;
    int nsplit = std::max(1, std::min(4096, 8192 / blocks));
    int64_t *ds = nullptr;
    double4 *part = nullptr, *out = nullptr;
    bool ok = check(hipMalloc(&ds, 8 * (size_t)ns), "hipMalloc") &&
              check(hipMalloc(&part, sizeof(double4) * (size_t)ns * nsplit), "hipMalloc") &&
              check(hipMalloc(&out, sizeof(double4) * (size_t)ns), "hipMalloc") &&
              check(hipMemcpy(ds, s.data(), 8 * (size_t)ns, hipMemcpyHostToDevice), "hipMemcpy") &&
              check(launch_direct64(P->v.bpos, P->tree.n, ds, ns, P->prm.soft_type, P->prm.eps,
                                    P->prm.G, part, nsplit, out,
                                    P->prm.periodic ? P->v.ewald64 : nullptr, P->prm.box,
                                    P->stream), "direct-sum kernel") &&
              check(hipMemcpyAsync(out4, out, sizeof(double4) * (size_t)ns, hipMemcpyDeviceToHost,
                                   P->stream), "hipMemcpyAsync") &&
              check(hipStreamSynchronize(P->stream), "hipStreamSynchronize");
    for (void *p : {(void *)ds, (void *)part, (void *)out})
        if (p)
            (void)hipFree(p);
    return ok ? 0 : -1;
}

int pgrav_plan_device_buffers(void *plan, void **out32, void **out64, void **bpos) {
    Plan *P = (Plan *)plan;
    if (!P) {
        fail("null plan");
        return -1;
    }
    if (out32)
        *out32 = P->v.out32;
    if (out64)
        *out64 = P->v.out64;
    if (bpos)
        *bpos = (void *)P->v.bpos;
    return 0;
}

int pgrav_plan_cells(void *plan, int64_t node0, int64_t count, double *cofm4, double *bmax,
                     double *raw35) {
    Plan *P = (Plan *)plan;
    clear_error();
    if (!P || node0 < 0 || count < 0 || node0 + count > P->nnodes) {
        fail("pgrav_plan_cells: bad node range");
        return -1;
    }
    try {
        P->ensure_moments_view();
        bool ok = check(hipStreamSynchronize(P->stream), "hipStreamSynchronize");
        if (ok && cofm4)
            ok = check(hipMemcpy(cofm4, P->mv.cpos + node0, sizeof(double4) * count, hipMemcpyDeviceToHost), "hipMemcpy");
        if (ok && bmax)
            ok = check(hipMemcpy(bmax, P->mv.bmax + node0, sizeof(double) * count, hipMemcpyDeviceToHost), "hipMemcpy");
        if (ok && raw35)
            ok = check(hipMemcpy(raw35, P->mv.raw + 35 * node0, sizeof(double) * 35 * count, hipMemcpyDeviceToHost), "hipMemcpy");
        return ok ? 0 : -1;
    } catch (const std::exception &e) {
        fail("pgrav_plan_cells: %s", e.what());
        return -1;
    }
}

int pgrav_plan_permutation(void *plan, int64_t *perm) {
    Plan *P = (Plan *)plan;
    if (!P || !perm) {
        fail("null argument");
        return -1;
    }
    memcpy(perm, P->tree.perm.data(), 8 * (size_t)P->tree.n);
    return 0;
}

}  // extern "C"
