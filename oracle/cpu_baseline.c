/*
 * TEST INFRASTRUCTURE ONLY -- CPU baseline timer.
 *
 * Times the reference's own CPU kernels (compiled from the reference sources
 * into oracle/_ref by oracle/Makefile) over interaction lists prepared by the
 * caller: for every sink of every sampled group, one call of the cell kernel
 * (e.g. do_gravph_sse4, nln/do_grav_sse4.c:1768) over the group's SoA-NSSE
 * cell blocks and one call of the body kernel (do_gravp_sse4,
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

typedef struct {
    grav_fn cell, body;
    int nsse, cell_sz;
    int ngroups;
    const int64_t *sink_off;
    const float *sinks3;
    const int64_t *cell_off, *body_off; /* in blocks */
    const float *cell_soa, *body_soa;
    float eps2;
    float *out4;
    atomic_int next;
} job_t;

static void *worker(void *arg) {
    job_t *J = (job_t *)arg;
    const size_t cblk = (size_t)J->nsse * J->cell_sz, bblk = (size_t)J->nsse * 4;
    for (;;) {
        int g = atomic_fetch_add(&J->next, 1);
        if (g >= J->ngroups)
            break;
        const float *c0 = J->cell_soa + cblk * J->cell_off[g];
        const float *c1 = J->cell_soa + cblk * J->cell_off[g + 1];
        const float *b0 = J->body_soa + bblk * J->body_off[g];
        const float *b1 = J->body_soa + bblk * J->body_off[g + 1];
        for (int64_t s = J->sink_off[g]; s < J->sink_off[g + 1]; s++) {
            float mass0 = 0, acc[3] = {0, 0, 0}, phi = 0, acc2[3] = {0, 0, 0}, phi2 = 0;
            int ncut = 0;
            if (c1 > c0)
                J->cell(c0, c1, J->sinks3 + 3 * s, &mass0, acc, &phi, &J->eps2, &ncut);
            if (b1 > b0)
                J->body(b0, b1, J->sinks3 + 3 * s, &mass0, acc2, &phi2, &J->eps2, &ncut);
            J->out4[4 * s + 0] = acc[0] + acc2[0];
            J->out4[4 * s + 1] = acc[1] + acc2[1];
            J->out4[4 * s + 2] = acc[2] + acc2[2];
            J->out4[4 * s + 3] = phi + phi2;
        }
    }
    return NULL;
}

static float *aligned_copy(const float *src, size_t n) {
    void *p = NULL;
    if (posix_memalign(&p, 64, (n ? n : 1) * sizeof(float)))
        return NULL;
    if (n)
        memcpy(p, src, n * sizeof(float));
    return (float *)p;
}

/* Returns wall seconds for `reps` passes over all groups, or -1 on error. */
double orc_cpu_ref_run(const char *lib_path, const char *cell_fn, const char *body_fn, int nsse,
                       int cell_sz, int ngroups, const int64_t *sink_off, const float *sinks3,
                       const int64_t *cell_off, const float *cell_soa, const int64_t *body_off,
                       const float *body_soa, float eps2, int nthreads, int reps, float *out4) {
    void *h = dlopen(lib_path, RTLD_NOW | RTLD_LOCAL);
    if (!h)
        return -1;
    grav_fn cf = (grav_fn)dlsym(h, cell_fn), bf = (grav_fn)dlsym(h, body_fn);
    if (!cf || !bf)
        return -1;
    const size_t nc = (size_t)cell_off[ngroups] * nsse * cell_sz;
    const size_t nb = (size_t)body_off[ngroups] * nsse * 4;
    float *cs = aligned_copy(cell_soa, nc), *bs = aligned_copy(body_soa, nb);
    if (!cs || !bs)
        return -1;
    if (nthreads < 1)
        nthreads = 1;
    if (nthreads > 256)
        nthreads = 256;
    pthread_t th[256];
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int r = 0; r < reps; r++) {
        job_t J = {cf, bf, nsse, cell_sz, ngroups, sink_off, sinks3, cell_off, body_off,
                   cs, bs, eps2, out4, 0};
        atomic_init(&J.next, 0);
        for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, worker, &J);
        for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    free(cs);
    free(bs);
    return (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
}
