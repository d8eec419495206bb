/*
 * TEST INFRASTRUCTURE ONLY -- CPU baseline timer.
 *
 * Times the reference's own CPU kernels (compiled from the reference sources
 * into oracle/_ref by oracle/Makefile) over the plan's interaction lists: for
 * every sink of every sampled group, one call of the cell kernel (e.g.
 * do_gravph_sse4, nln/do_grav_sse4.c:1768) over the group's SoA-NSSE cell
 * blocks and one call of the body kernel (do_gravp_sse4,
 * nln/do_grav_sse4.c:2167) over its body blocks -- the leaf evaluation of
 * nln/mac.c:477-630.  Groups are spread over `nthreads` pthreads (the
 * reference runs one single-threaded process per core; threads over disjoint
 * groups are the same work distribution).
 */
#include <dlfcn.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef void (*grav_fn)(const float *f, const float *fend, const float *pos0, float *mass0,
                        float *acc, float *phi0, const float *eps2p, int *ncut);

/* The caller passes the record table, the sorted
 * sinks and bodies and the CSR lists; each thread gathers one group at a
 * time into SoA-NSSE blocks (as nln/mac.c's append macros fill Hvec/Mvec)
 * and then runs the reference kernels for every sink of the group.  Only the
 * kernel calls are timed (kernel_seconds, summed over threads).
 */
typedef struct {
    grav_fn cell, body;
    int nsse, csz;
    const float *rec, *sinks3, *bodies4;
    const int64_t *gfirst;
    const int *gcount;
    const int64_t *hoff, *boff;
    const uint32_t *hidx, *bidx;
    const int *groups;
    int ng;
    float eps2;
    float *out4;
    atomic_int next;
    double ksec[256];
} ljob_t;

typedef struct {
    ljob_t *J;
    int tid;
} larg_t;

static int grow(float **buf, size_t *cap, size_t need) {
    if (need <= *cap)
        return 1;
    free(*buf);
    *cap = need + need / 2 + 64;
    void *p = NULL;
    if (posix_memalign(&p, 64, *cap * sizeof(float)))
        return 0;
    *buf = (float *)p;
    return 1;
}

static void *lworker(void *arg) {
    ljob_t *J = ((larg_t *)arg)->J;
    const int tid = ((larg_t *)arg)->tid;
    const int ns = J->nsse, csz = J->csz;
    float *cb = NULL, *bb = NULL;
    size_t ccap = 0, bcap = 0;
    double ks = 0;
    for (;;) {
        const int k = atomic_fetch_add(&J->next, 1);
        if (k >= J->ng)
            break;
        const int g = J->groups[k];
        const int64_t h0 = J->hoff[g], nh = J->hoff[g + 1] - h0;
        const int64_t b0 = J->boff[g], nb = J->boff[g + 1] - b0;
        const int64_t nhb = (nh + ns - 1) / ns, nbb = (nb + ns - 1) / ns;
        if (!grow(&cb, &ccap, (size_t)(nhb * ns * csz)) || !grow(&bb, &bcap, (size_t)(nbb * ns * 4)))
            break;
        memset(cb, 0, sizeof(float) * nhb * ns * csz);
        memset(bb, 0, sizeof(float) * nbb * ns * 4);
        for (int64_t i = 0; i < nh; i++) {
            const float *r = J->rec + (size_t)J->hidx[h0 + i] * csz;
            float *d = cb + (i / ns) * ns * csz + (i % ns);
            for (int f = 0; f < csz; f++) d[f * ns] = r[f];
        }
        for (int64_t i = 0; i < nb; i++) {
            const float *r = J->bodies4 + (size_t)J->bidx[b0 + i] * 4;
            float *d = bb + (i / ns) * ns * 4 + (i % ns);
            for (int f = 0; f < 4; f++) d[f * ns] = r[f];
        }
        struct timespec t0, t1;
        clock_gettime(CLOCK_MONOTONIC, &t0);
        for (int64_t s = J->gfirst[g]; s < J->gfirst[g] + J->gcount[g]; s++) {
            float mass0 = 0, acc[3] = {0, 0, 0}, phi = 0, acc2[3] = {0, 0, 0}, phi2 = 0;
            int ncut = 0;
            if (nhb)
                J->cell(cb, cb + nhb * ns * csz, J->sinks3 + 3 * s, &mass0, acc, &phi, &J->eps2, &ncut);
            if (nbb)
                J->body(bb, bb + nbb * ns * 4, J->sinks3 + 3 * s, &mass0, acc2, &phi2, &J->eps2, &ncut);
            J->out4[4 * s + 0] = acc[0] + acc2[0];
            J->out4[4 * s + 1] = acc[1] + acc2[1];
            J->out4[4 * s + 2] = acc[2] + acc2[2];
            J->out4[4 * s + 3] = phi + phi2;
        }
        clock_gettime(CLOCK_MONOTONIC, &t1);
        ks += (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
    }
    J->ksec[tid] = ks;
    free(cb);
    free(bb);
    return NULL;
}

/* Returns wall seconds of the parallel region (gathers included), or -1 on
 * error; *kernel_seconds = kernel-call seconds summed over the threads. */
double orc_cpu_ref_run_lists(const char *lib_path, const char *cell_fn, const char *body_fn,
                             int nsse, int cell_sz, const float *rec, const float *sinks3,
                             const float *bodies4, const int64_t *gfirst, const int *gcount,
                             const int64_t *hoff, const uint32_t *hidx, const int64_t *boff,
                             const uint32_t *bidx, const int *groups, int ng, float eps2,
                             int nthreads, float *out4, double *kernel_seconds) {
    void *h = dlopen(lib_path, RTLD_NOW | RTLD_LOCAL);
    if (!h)
        return -1;
    grav_fn cf = (grav_fn)dlsym(h, cell_fn), bf = (grav_fn)dlsym(h, body_fn);
    if (!cf || !bf)
        return -1;
    if (nthreads < 1)
        nthreads = 1;
    if (nthreads > 256)
        nthreads = 256;
    static ljob_t J;
    memset(&J, 0, sizeof(J));
    J.cell = cf;
    J.body = bf;
    J.nsse = nsse;
    J.csz = cell_sz;
    J.rec = rec;
    J.sinks3 = sinks3;
    J.bodies4 = bodies4;
    J.gfirst = gfirst;
    J.gcount = gcount;
    J.hoff = hoff;
    J.hidx = hidx;
    J.boff = boff;
    J.bidx = bidx;
    J.groups = groups;
    J.ng = ng;
    J.eps2 = eps2;
    J.out4 = out4;
    atomic_init(&J.next, 0);
    pthread_t th[256];
    larg_t args[256];
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int t = 0; t < nthreads; t++) {
        args[t].J = &J;
        args[t].tid = t;
        pthread_create(&th[t], NULL, lworker, &args[t]);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    double ks = 0;
    for (int t = 0; t < nthreads; t++) ks += J.ksec[t];
    if (kernel_seconds)
        *kernel_seconds = ks;
    return (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
}
