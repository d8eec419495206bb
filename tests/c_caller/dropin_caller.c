/* A C host that drives the drop-in ABI the way nln/mac.c does
 * (nln/mac.c:265-277, 964-971, 1060-1066): WalkInitSinkCUDA on the body table,
 * per sink group one slot from qallocCUDA(NULL), the group's order-4 cells
 * through grav_mn_CUDA("pH") and its source bodies through
 * grav_mns_CUDA("pMM_sK1"), then WalkTerminateSinkCUDA adds the results into
 * the table.  Built by the tests with a C compiler against include/pgrav.h and
 * libpgrav.so only.
 *
 * usage: dropin_caller IN OUT
 *   IN : int32 n, ngroups, source_n, nseg; float eps_inv, mmass;
 *        float btab[14 n]; int32 groups[2 ngroups] (first body, count);
 *        float hvec[29 source_n] (SoA-8 blocks); uint32 seg[2 nseg] (base, length)
 *   OUT: float btab[14 n] after WalkTerminateSinkCUDA
 */
#include <stdio.h>
#include <stdlib.h>

#include "pgrav.h"

static void *slurp(FILE *f, size_t bytes) {
    void *p = malloc(bytes ? bytes : 1);
    if (!p || fread(p, 1, bytes, f) != bytes) {
        fprintf(stderr, "short input\n");
        exit(2);
    }
    return p;
}

int main(int argc, char **argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: %s IN OUT\n", argv[0]);
        return 2;
    }
    FILE *in = fopen(argv[1], "rb");
    if (!in) {
        perror(argv[1]);
        return 2;
    }
    int32_t hdr[4];
    float par[2];
    if (fread(hdr, sizeof hdr, 1, in) != 1 || fread(par, sizeof par, 1, in) != 1) {
        fprintf(stderr, "short header\n");
        return 2;
    }
    const int n = hdr[0], ngroups = hdr[1], source_n = hdr[2], nseg = hdr[3];
    float *btab = slurp(in, sizeof(float) * 14 * (size_t)n);
    int32_t *groups = slurp(in, sizeof(int32_t) * 2 * (size_t)ngroups);
    float *hvec = slurp(in, sizeof(float) * 29 * (size_t)source_n);
    segment *seg = slurp(in, sizeof(segment) * (size_t)nseg);
    fclose(in);
    int bodies = 0;
    for (int s = 0; s < nseg; s++) bodies += (int)seg[s].length;

    CUDA_Init(NULL);
    if (vecwidthCUDA() != 4) {
        fprintf(stderr, "vecwidthCUDA\n");
        return 1;
    }
    WalkInitSinkCUDA(btab, 14, n);
    for (int g = 0; g < ngroups; g++) {
        const int first = groups[2 * g], m = groups[2 * g + 1];
        int q;
        while ((q = qallocCUDA(NULL)) < 0) {
        }
        grav_mn_CUDA("pH", btab + 14 * (size_t)first, NULL, m, 14, hvec, source_n, 29, 0.0f,
                     NULL, q);
        grav_mns_CUDA("pMM_sK1", first, m, seg, nseg, bodies, par[1], par[0], NULL, q);
    }
    WalkTerminateSinkCUDA(btab, 14, n);
    if (pgrav_last_error()) {
        fprintf(stderr, "%s\n", pgrav_last_error());
        return 1;
    }
    FILE *out = fopen(argv[2], "wb");
    if (!out || fwrite(btab, sizeof(float) * 14, (size_t)n, out) != (size_t)n) {
        perror(argv[2]);
        return 2;
    }
    fclose(out);
    free(btab);
    free(groups);
    free(hvec);
    free(seg);
    puts("ok");
    return 0;
}
