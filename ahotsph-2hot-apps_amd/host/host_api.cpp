// Host-only view of the tree and interaction lists (no device needed): lets a
// caller or a test inspect exactly what the batched kernel will evaluate.
#include <string.h>

#include <new>
#include <stdexcept>

#include "pgrav.h"
#include "pgrav_host.h"
#include "pgrav_record.h"

namespace pgrav {
void fail(const char *fmt, ...);
void clear_error();
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
    if (!pos3 || !mass || n <= 0 || !p || !(p->theta > 0) || p->group_max < 1 ||
        p->group_max > 64 || (p->order != 0 && p->order != 2 && p->order != 4)) {
        fail("pgrav_host_plan: bad arguments");
        return nullptr;
    }
    HostPlan *H = new (std::nothrow) HostPlan();
    if (!H) {
        fail("out of host memory");
        return nullptr;
    }
    try {
        H->tp.theta = p->theta;
        H->tp.group_max = p->group_max;
        H->tp.leaf_max = p->leaf_max > 0 ? p->leaf_max : 16;
        H->tp.order = p->order;
        H->tp.pcut = p->pcut;
        H->tp.periodic = p->periodic;
        H->tp.nimage = p->nimage;
        if (p->rcb_cell > 0)
            H->tp.rcb_cell = p->rcb_cell;
        H->tp.box = p->box;
        H->tp.nthreads = p->nthreads;
        H->tp.merge_groups = p->merge_groups;
        H->tp.dispatch = p->dispatch;
        H->tp.mixed_order = p->mixed_order;
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
