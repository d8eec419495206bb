/*
 * libpgrav -- MI355X (gfx950) implementation of 2HOT's gravity interaction-list
 * evaluation.  C ABI: plain pointers and sizes only.
 *
 * Part 1 is a drop-in for the reference's nln/gravcuda.h: same symbol names,
 * argument meaning, ownership (caller owns every host array; each grav_* call
 * copies its list before returning) and accumulation semantics (results are
 * added into the body table by WalkTerminateSinkCUDA).  Part 2 mirrors the
 * Python-facing C routines of nln/py_pgrav and nln/py_pgrav_cu.  Part 3 is the
 * batched, mixed-precision path (no reference counterpart; see each comment).
 *
 * Errors: like the reference's Error(), any runtime error aborts the process
 * by default.  pgrav_set_error_mode(1) makes every entry point record the
 * message and return instead (pgrav_last_error()).
 */
#ifndef PGRAV_H
#define PGRAV_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* nln/segment.h:16-19 */
typedef struct {
    unsigned int base;
    unsigned int length;
} segment;

/* ------------------------------------------------------------------ part 1
 * nln/gravcuda.h:13-69 */

/* nln/gravcuda.h:14 (impl nln/cuda.cu:13): select device 0 and create the
 * library state.  device_prop is ignored (the reference's definition takes no
 * argument). */
void CUDA_Init(void *device_prop);
/* nln/gravcuda.h:15 (declared, never defined in the reference): releases every
 * device and pinned buffer the library holds. */
void CUDA_Finalize(void);
/* nln/gravcuda.h:16 (impl nln/pgrav.cu:1022): returns 4, the reference's
 * VECWIDTH; callers only print it (nln/main.c:688). */
int vecwidthCUDA(void);
/* nln/gravcuda.h:17 (impl nln/pgrav.cu:1024, 814-856): returns a free slot
 * index in [0,32) or -1 when all 32 slots still have work in flight; *inuse
 * (may be NULL) receives the number of busy slots. */
int qallocCUDA(int *inuse);
/* nln/gravcuda.h:18 (impl nln/pgrav.cu:878-906): uploads the cell table for
 * grav_qns_CUDA.  qtab points at nobj hexacell records (nln/physics.h:140-157)
 * and stride is their size in BYTES (what nln/mac.c:257 passes). */
void WalkInitSrcCUDA(float *qtab, int stride, int64_t nobj);
/* nln/gravcuda.h:19 (impl nln/pgrav.cu:908-912) */
void WalkTerminateSrcCUDA(void);
/* nln/gravcuda.h:20 (impl nln/pgrav.cu:914-967): btab is the body table
 * (stride floats per body, position at +1..3); copies positions to the device
 * and zeroes the device accumulator. */
void WalkInitSinkCUDA(float *btab, int stride, int64_t nobj);
/* nln/gravcuda.h:21 (impl nln/pgrav.cu:969-1020): waits for every slot, then
 * btab[i*stride+7..10] += (ax, ay, az, phi) accumulated since WalkInitSinkCUDA. */
void WalkTerminateSinkCUDA(float *btab, int stride, int64_t nobj);
/* nln/gravcuda.h:22-32 (impl nln/pgrav.cu:1026-1087): routine "pM", "pQ",
 * "pH" or "pM_sK1"; p points into the body table passed to WalkInitSinkCUDA
 * (its offset selects n sinks); f is an SoA-8 source list of source_n records
 * of sz floats; e = 1/eps for pM_sK1.  accp and ncut are unused, as in the
 * reference.  An unknown routine name is ignored, as in the reference. */
void grav_mn_CUDA(const char *routine, const float *p, float *accp, const int n,
                  const int stride, const float *f, const int source_n, const int sz,
                  const float e, int *ncut, int q);
/* nln/gravcuda.h:33-42 (impl nln/pgrav.cu:1089-1145): routine "pMM_sK1";
 * sinks base..base+m-1 against the bodies named by seg[0..seg_n). */
void grav_mns_CUDA(const char *routine, const int base, const int m, const segment *seg,
                   const int seg_n, const int source_n, const float mmass, float e, int *ncut,
                   int q);
/* nln/gravcuda.h:43-56 (impl nln/pgrav.cu:1202-1295): routine "pMM_mnss_sK1";
 * sink i of sink_base..+m uses seg[ss_seg[ss_index[i]]]. */
void grav_mnss_CUDA(const char *routine, const int sink_base, const int m,
                    const uint16_t *ss_index, const segment *ss_seg, const int ss_len,
                    const segment *seg, const int seg_len, const int source_base,
                    const int *source_n, const float mmass, float e, int *ncut, int q);
/* nln/gravcuda.h:57-62 (impl nln/pgrav.cu:1147-1200): routine "pQQ1" (also
 * accepts "pQ1", the name the only caller nln/mac.c:1143 passes); sinks
 * base..base+m-1 against cells source_list[] of the WalkInitSrcCUDA table. */
void grav_qns_CUDA(const char *routine, const int base, const int m, const int *source_list,
                   const int source_n, int q);
/* nln/gravcuda.h:63-69 (impl nln/pgrav.cu:1297-1358): routine "pQ1_qnss";
 * sink i uses run seg[i] of the packed SoA-8 order-2 buffer `source`. */
void grav_qnss_CUDA(const char *routine, const int base, const int m, const segment *seg,
                    const float *source, const int source_n, int q);

/* ------------------------------------------------------------------ part 2 */

/* nln/py_pgrav_cu/pgrav.cu:510-564: grav_mn_CUDA dispatched by record size
 * (sz 4 -> pM, 13 -> pQ, 29 -> pH). */
void pinteractCUDA(const float *p, float *accp, const int m, const int stride, const float *f,
                   const int source_n, const int sz, int q);

/* nln/py_pgrav/pgrav.pyx:21-47 (monopole / vec_monopole, which call
 * _pgravm and pMinteract, nln/pgrav_vec.c:259-300): self-contained M x N on
 * the GPU from host arrays.  Sink i has position p[i*stride+1..3];
 * accp[i*stride+0..3] += (ax, ay, az, phi).  kind: 0 pM, 2 pQ, 4 pH, 5 pM_sK1
 * (e = 1/eps).  layout 0: f holds source_n AoS records of sz floats;
 * layout 1: SoA-8 blocks (source_n padded to a multiple of 8).
 * Returns 0 or -1 on error. */
int pgrav_mn_host(int kind, const float *p, float *accp, int m, int stride, const float *f,
                  int source_n, int sz, int layout, float e);

/* ------------------------------------------------------------------ control
 * (no reference counterpart) */
int pgrav_set_error_mode(int mode);   /* 0 abort (reference), 1 record and return */
const char *pgrav_last_error(void);   /* NULL when no error is recorded */
void pgrav_clear_error(void);
/* +1 (default): phi of the quadrupole term with the physically consistent sign
 * of nln/do_grav_sse4.c:1836 and nln/pgrav_vec.c:117; -1: the sign of
 * nln/pgrav.cu:114,277 ("phi -= eqe").  Accelerations are identical. */
void pgrav_set_quad_phi_sign(int sign);
int pgrav_synchronize(void);

#ifdef __cplusplus
}
#endif
#endif
