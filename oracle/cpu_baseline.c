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
#define _GNU_SOURCE
#include <dlfcn.h>
#include <pthread.h>
#include <sched.h>
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
    double ksec[256][2];  /* per thread: seconds in cell calls, in body calls */
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
    double kc = 0, kb = 0;
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
        /* the cell kernel for every sink of the group, then the body kernel, each
         * timed as one block (the reference's per-kind timers, nln/main.c:1259-1275) */
        const int64_t s0 = J->gfirst[g], s1 = s0 + J->gcount[g];
        struct timespec t0, t1, t2;
        clock_gettime(CLOCK_MONOTONIC, &t0);
        for (int64_t s = s0; s < s1; s++) {
            float mass0 = 0, acc[3] = {0, 0, 0}, phi = 0;
            int ncut = 0;
            if (nhb)
                J->cell(cb, cb + nhb * ns * csz, J->sinks3 + 3 * s, &mass0, acc, &phi, &J->eps2, &ncut);
            J->out4[4 * s + 0] = acc[0];
            J->out4[4 * s + 1] = acc[1];
            J->out4[4 * s + 2] = acc[2];
            J->out4[4 * s + 3] = phi;
        }
        clock_gettime(CLOCK_MONOTONIC, &t1);
        for (int64_t s = s0; s < s1; s++) {
            float mass0 = 0, acc2[3] = {0, 0, 0}, phi2 = 0;
            int ncut = 0;
            if (nbb)
                J->body(bb, bb + nbb * ns * 4, J->sinks3 + 3 * s, &mass0, acc2, &phi2, &J->eps2, &ncut);
            J->out4[4 * s + 0] += acc2[0];
            J->out4[4 * s + 1] += acc2[1];
            J->out4[4 * s + 2] += acc2[2];
            J->out4[4 * s + 3] += phi2;
        }
        clock_gettime(CLOCK_MONOTONIC, &t2);
        kc += (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
        kb += (t2.tv_sec - t1.tv_sec) + 1e-9 * (t2.tv_nsec - t1.tv_nsec);
    }
    J->ksec[tid][0] = kc;
    J->ksec[tid][1] = kb;
    free(cb);
    free(bb);
    return NULL;
}

/* Returns wall seconds of the parallel region (gathers included), or -1 on
 * error; kernel_seconds[0], [1] = seconds in cell-kernel and body-kernel calls,
 * each summed over the threads.  cpus (may be NULL): thread t runs pinned to
 * logical CPU cpus[t] (one per physical core, as one reference process per
 * core would be). */
double orc_cpu_ref_run_lists(const char *lib_path, const char *cell_fn, const char *body_fn,
                             int nsse, int cell_sz, const float *rec, const float *sinks3,
                             const float *bodies4, const int64_t *gfirst, const int *gcount,
                             const int64_t *hoff, const uint32_t *hidx, const int64_t *boff,
                             const uint32_t *bidx, const int *groups, int ng, float eps2,
                             int nthreads, const int *cpus, float *out4,
                             double *kernel_seconds) {
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
    int started = 0;
    for (int t = 0; t < nthreads; t++) {
        args[t].J = &J;
        args[t].tid = t;
        pthread_attr_t at;
        pthread_attr_init(&at);
        if (cpus) {
            cpu_set_t set;
            CPU_ZERO(&set);
            CPU_SET(cpus[t], &set);
            pthread_attr_setaffinity_np(&at, sizeof(set), &set);
        }
        if (pthread_create(&th[t], &at, lworker, &args[t]) == 0)
            started++;
        pthread_attr_destroy(&at);
        if (started != t + 1)
            break;
    }
    for (int t = 0; t < started; t++) pthread_join(th[t], NULL);
    if (started != nthreads)
        return -1;
    clock_gettime(CLOCK_MONOTONIC, &t1);
    double kc = 0, kb = 0;
    for (int t = 0; t < nthreads; t++) {
        kc += J.ksec[t][0];
        kb += J.ksec[t][1];
    }
    if (kernel_seconds) {
        kernel_seconds[0] = kc;
        kernel_seconds[1] = kb;
    }
    return (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
}
