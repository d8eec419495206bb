// Host-only view of the tree and interaction lists (no device needed): lets a
// caller or a test inspect exactly what the batched kernel will evaluate.
#include <stdio.h>
#include <string.h>

#include <new>
#include <stdexcept>

#include "pgrav.h"
#include "pgrav_host.h"
#include "pgrav_record.h"

namespace pgrav {
void fail(const char *fmt, ...);
void clear_error();

std::string tree_params_from(const pgrav_params *p, TreeParams &tp) {
    char msg[256];
    msg[0] = 0;
    if (p->order != 0 && p->order != 2 && p->order != 4)
        snprintf(msg, sizeof msg, "order must be 0, 2 or 4 (got %d)", p->order);
    else if (p->group_max < 1 || p->group_max > 64)
        snprintf(msg, sizeof msg, "group_max must be 1..64 (got %d)", p->group_max);
    else if (p->mac != 0 && p->mac != 1)
        snprintf(msg, sizeof msg, "mac must be 0 (opening angle) or 1 (error tolerance)");
    else if (p->mac == 0 && !(p->theta > 0))
        snprintf(msg, sizeof msg, "theta must be > 0 (got %g)", p->theta);
    else if (p->mac == 1 && (!(p->errtol > 0) || !(p->frac_tol >= 0) || !(p->frac_tol0 >= 0) ||
                             !(p->ptol_boost >= 0) || p->quad_ncut < 0 || p->hexa_ncut < 0))
        snprintf(msg, sizeof msg,
                 "mac 1 needs errtol > 0 and frac_tol, frac_tol0, ptol_boost, quad_ncut, "
                 "hexa_ncut >= 0 (errtol %g frac_tol %g frac_tol0 %g)",
                 p->errtol, p->frac_tol, p->frac_tol0);
    else if (p->merge_groups < 0 || p->merge_groups > 5)
        snprintf(msg, sizeof msg, "merge_groups must be 0..5 (got %d)", p->merge_groups);
    else if (p->dispatch < 0 || p->dispatch > 3)
        snprintf(msg, sizeof msg, "dispatch must be 0..3 (got %d)", p->dispatch);
    else if (p->leaf_max < 0 || p->pcut < 0 || p->rcb_cell < 0 || p->nthreads < 0)
        snprintf(msg, sizeof msg, "leaf_max, pcut, rcb_cell and nthreads must be >= 0");
    else if (p->mixed_order != 0 && p->mixed_order != 1)
        snprintf(msg, sizeof msg, "mixed_order must be 0 or 1 (got %d)", p->mixed_order);
    else if (p->group_box != 0 && p->group_box != 1)
        snprintf(msg, sizeof msg, "group_box must be 0 (sphere) or 1 (bounding box)");
    else if (p->periodic < 0 || p->periodic > 2 || (p->periodic && !(p->box > 0)))
        snprintf(msg, sizeof msg, "periodic must be 0..2 with box > 0 (periodic %d box %g)",
                 p->periodic, p->box);
    else if (p->periodic == 2 && (p->nimage < 0 || p->nimage > MAX_NIMAGE))
        snprintf(msg, sizeof msg, "nimage must be 0..%d (got %d)", MAX_NIMAGE, p->nimage);
    if (msg[0])
        return msg;
    tp.theta = p->theta;
    tp.group_max = p->group_max;
    tp.leaf_max = p->leaf_max > 0 ? p->leaf_max : 16;
    tp.order = p->order;
    tp.pcut = p->pcut;
    tp.periodic = p->periodic;
    tp.nimage = p->nimage;
    if (p->rcb_cell > 0)
        tp.rcb_cell = p->rcb_cell;
    tp.box = p->box;
    tp.nthreads = p->nthreads;
    tp.merge_groups = p->merge_groups;
    tp.dispatch = p->dispatch;
    tp.mixed_order = p->mixed_order;
    tp.mac = p->mac;
    tp.errtol = p->errtol;
    tp.frac_tol = p->frac_tol;
    tp.frac_tol0 = p->frac_tol0;
    tp.ptol_boost = p->ptol_boost;
    tp.stol_max = p->stol_max;
    tp.quad_ncut = p->quad_ncut;
    tp.hexa_ncut = p->hexa_ncut;
    tp.group_box = p->group_box;
    return std::string();
}
}  // namespace pgrav

using namespace pgrav;

namespace {
struct HostPlan {
    TreeParams tp;
    Tree tree;
    Lists lists;
};
}  // namespace

extern "C" {

void *pgrav_host_plan(const double *pos3, const double *mass, int64_t n, const pgrav_params *p) {
    clear_error();
    if (!pos3 || !mass || n <= 0 || !p) {
        fail("pgrav_host_plan: bad arguments");
        return nullptr;
    }
    HostPlan *H = new (std::nothrow) HostPlan();
    if (!H) {
        fail("out of host memory");
        return nullptr;
    }
    const std::string bad = tree_params_from(p, H->tp);
    if (!bad.empty()) {
        fail("pgrav_host_plan: %s", bad.c_str());
        delete H;
        return nullptr;
    }
    try {
        build_tree(pos3, mass, n, H->tp, H->tree);
        build_groups(H->tree, H->tp, H->lists);
        build_lists(H->tree, H->tp, H->lists);
    } catch (const std::exception &e) {
        fail("pgrav_host_plan: %s", e.what());
        delete H;
        return nullptr;
    }
    return H;
}

void pgrav_host_destroy(void *h) { delete (HostPlan *)h; }

int pgrav_host_stats(void *h, pgrav_stats *st) {
    HostPlan *H = (HostPlan *)h;
    if (!H || !st) {
        fail("null argument");
        return -1;
    }
    memset(st, 0, sizeof(*st));
    st->n = H->tree.n;
    st->nnodes = H->tree.nnodes();
    st->ngroups = H->lists.ngroups;
    for (int k = 0; k < NLIST; k++) {
        st->pairs[k] = H->lists.pairs[k];
        st->entries[k] = H->lists.off[k].back();
    }
    st->mean_group = H->lists.ngroups ? (double)H->tree.n / H->lists.ngroups : 0.0;
    return 0;
}

int pgrav_host_group(void *h, int64_t g, int64_t *first, int *count, double *center4) {
    HostPlan *H = (HostPlan *)h;
    if (!H || g < 0 || g >= H->lists.ngroups) {
        fail("bad group");
        return -1;
    }
    if (first)
        *first = H->lists.gfirst[g];
    if (count)
        *count = H->lists.gcount[g];
    if (center4)
        memcpy(center4, &H->lists.gcenter[4 * g], 4 * sizeof(double));
    return 0;
}

int64_t pgrav_host_list(void *h, int64_t g, int kind, uint32_t *out, int64_t cap) {
    HostPlan *H = (HostPlan *)h;
    if (!H || g < 0 || g >= H->lists.ngroups || kind < 0 || kind >= NLIST) {
        fail("bad group or list kind");
        return -1;
    }
    const int64_t a = H->lists.off[kind][g], b = H->lists.off[kind][g + 1];
    if (out)
        for (int64_t i = 0; i < b - a && i < cap; i++) out[i] = H->lists.idx[kind][a + i];
    return b - a;
}

int64_t pgrav_host_group_slots(void *h, int64_t g, int64_t *slots) {
    HostPlan *H = (HostPlan *)h;
    if (!H || g < 0 || g >= H->lists.ngroups) {
        fail("bad group");
        return -1;
    }
    const Lists &L = H->lists;
    const int64_t n = L.gcount[g];
    if (slots)
        for (int64_t i = 0; i < n; i++)
            slots[i] = L.gfirst[g] >= 0 ? L.gfirst[g] + i : L.gslots[L.gslot_off[g] + i];
    return n;
}

int pgrav_host_group_slot_table(void *h, int64_t *off, int64_t *slots) {
    HostPlan *H = (HostPlan *)h;
    if (!H || !off || !slots) {
        fail("pgrav_host_group_slot_table: null argument");
        return -1;
    }
    const Lists &L = H->lists;
    off[0] = 0;
    for (int64_t g = 0; g < L.ngroups; g++) {
        const int64_t n = L.gcount[g];
        for (int64_t i = 0; i < n; i++)
            slots[off[g] + i] = L.gfirst[g] >= 0 ? L.gfirst[g] + i : L.gslots[L.gslot_off[g] + i];
        off[g + 1] = off[g] + n;
    }
    return 0;
}

int64_t pgrav_host_list_images(void *h, int64_t g, int kind, uint8_t *out, int64_t cap) {
    HostPlan *H = (HostPlan *)h;
    if (!H || g < 0 || g >= H->lists.ngroups || kind < 0 || kind >= NLIST) {
        fail("bad group or list kind");
        return -1;
    }
    if (H->lists.img[kind].empty() && H->lists.off[kind][H->lists.ngroups] > 0) {
        fail("the lists carry no images (periodic != 2)");
        return -1;
    }
    const int64_t a = H->lists.off[kind][g], b = H->lists.off[kind][g + 1];
    if (out)
        for (int64_t i = 0; i < b - a && i < cap; i++) out[i] = H->lists.img[kind][a + i];
    return b - a;
}

int pgrav_host_groups(void *h, int64_t *first, int *count, double *center4) {
    HostPlan *H = (HostPlan *)h;
    if (!H) {
        fail("null plan");
        return -1;
    }
    const Lists &L = H->lists;
    if (first)
        memcpy(first, L.gfirst.data(), sizeof(int64_t) * L.ngroups);
    if (count)
        memcpy(count, L.gcount.data(), sizeof(int) * L.ngroups);
    if (center4)
        memcpy(center4, L.gcenter.data(), sizeof(double) * 4 * L.ngroups);
    return 0;
}

int64_t pgrav_host_csr(void *h, int kind, int64_t *off, uint32_t *idx) {
    HostPlan *H = (HostPlan *)h;
    if (!H || kind < 0 || kind >= NLIST) {
        fail("bad list kind");
        return -1;
    }
    const Lists &L = H->lists;
    if (off)
        memcpy(off, L.off[kind].data(), sizeof(int64_t) * (L.ngroups + 1));
    if (idx)
        memcpy(idx, L.idx[kind].data(), sizeof(uint32_t) * L.idx[kind].size());
    return (int64_t)L.idx[kind].size();
}

int pgrav_host_cell(void *h, int64_t node, double *cofm3, double *bmax, double *raw35,
                    int64_t *first, int *count, int *child, int *nchild) {
    HostPlan *H = (HostPlan *)h;
    if (!H || node < 0 || node >= H->tree.nnodes()) {
        fail("bad node");
        return -1;
    }
    const Tree &t = H->tree;
    if (cofm3)
        memcpy(cofm3, &t.cofm[3 * node], 3 * sizeof(double));
    if (bmax)
        *bmax = t.bmax[node];
    if (raw35)
        memcpy(raw35, &t.raw[35 * node], 35 * sizeof(double));
    if (first)
        *first = t.first[node];
    if (count)
        *count = t.count[node];
    if (child)
        *child = t.child[node];
    if (nchild)
        *nchild = t.nchild[node];
    return 0;
}

int pgrav_host_rcrit(void *h, double *rcrit3) {
    HostPlan *H = (HostPlan *)h;
    if (!H || !rcrit3) {
        fail("pgrav_host_rcrit: null argument");
        return -1;
    }
    if (H->tp.mac != 1) {
        fail("pgrav_host_rcrit: the plan uses the opening-angle MAC (no critical radii)");
        return -1;
    }
    memcpy(rcrit3, H->tree.rcrit.data(), sizeof(double) * H->tree.rcrit.size());
    return 0;
}

int pgrav_host_arel_rcrit(double m, double bmax, double B2, double B4, double B6, int64_t count,
                          double r0, const pgrav_params *p, double *rcrit3) {
    clear_error();
    if (!p || !rcrit3 || !(r0 > 0)) {
        fail("pgrav_host_arel_rcrit: bad arguments");
        return -1;
    }
    TreeParams tp;
    pgrav_params q = *p;
    q.mac = 1;
    const std::string bad = tree_params_from(&q, tp);
    if (!bad.empty()) {
        fail("pgrav_host_arel_rcrit: %s", bad.c_str());
        return -1;
    }
    arel_rcrit(m, bmax, B2, B4, B6, count, tp, r0, rcrit3);
    return 0;
}

int pgrav_host_record(void *h, int64_t node, double *rec29) {
    HostPlan *H = (HostPlan *)h;
    if (!H || node < 0 || node >= H->tree.nnodes() || !rec29) {
        fail("bad node");
        return -1;
    }
    const Tree &t = H->tree;
    const double bm = t.bmax[node] > 0 ? t.bmax[node] : 1.0;
    raw_to_record<double>(&t.raw[35 * node], &t.cofm[3 * node], bm, rec29);
    rec29[5] = rec29[6] = rec29[7] = 0.0;  // zero dipole about the centre of mass
    return 0;
}

int pgrav_host_records(void *h, const uint32_t *nodes, int64_t n, double *rec29) {
    for (int64_t i = 0; i < n; i++)
        if (pgrav_host_record(h, nodes[i], rec29 + 29 * i))
            return -1;
    return 0;
}

int pgrav_host_sorted(void *h, int64_t *perm, double *pos3) {
    HostPlan *H = (HostPlan *)h;
    if (!H) {
        fail("null plan");
        return -1;
    }
    if (perm)
        memcpy(perm, H->tree.perm.data(), 8 * (size_t)H->tree.n);
    if (pos3)
        memcpy(pos3, H->tree.pos.data(), 24 * (size_t)H->tree.n);
    return 0;
}

int pgrav_host_ewald_le(const double *pos3, const double *mass, int64_t n, double box, int nimage,
                        double G, double *acc4) {
    clear_error();
    if (!pos3 || !mass || !acc4 || n < 0 || !(box > 0) || nimage < 0 || nimage > MAX_NIMAGE) {
        fail("pgrav_host_ewald_le: bad arguments");
        return -1;
    }
    try {
        le_evaluate_host(pos3, mass, n, box, nimage, G, acc4);
    } catch (const std::exception &e) {
        fail("pgrav_host_ewald_le: %s", e.what());
        return -1;
    }
    return 0;
}

int pgrav_host_ewald_le_sb(const double *pos3, const double *mass, int64_t n, double box,
                           int nimage, double G, double *acc4) {
    clear_error();
    if (!pos3 || !mass || !acc4 || n < 0 || !(box > 0) || nimage < 0 || nimage > MAX_NIMAGE) {
        fail("pgrav_host_ewald_le_sb: bad arguments");
        return -1;
    }
    try {
        le_evaluate_host(pos3, mass, n, box, nimage, G, acc4, true);
    } catch (const std::exception &e) {
        fail("pgrav_host_ewald_le_sb: %s", e.what());
        return -1;
    }
    return 0;
}

int pgrav_host_le_lattice(int k, int nimage, double *ck) {
    clear_error();
    if (!ck || k < 4 || k > 8 || (k & 1) || nimage < 0) {
        fail("pgrav_host_le_lattice: k in {4, 6, 8}, nimage >= 0");
        return -1;
    }
    *ck = le_lattice_coeff(k, nimage);
    return 0;
}

int pgrav_host_lpt(void *h, int32_t *order) {
    HostPlan *H = (HostPlan *)h;
    if (!H || !order) {
        fail("null argument");
        return -1;
    }
    memcpy(order, H->lists.lpt.data(), 4 * (size_t)H->lists.ngroups);
    return 0;
}

}  // extern "C"
