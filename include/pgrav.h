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
 * (may be NULL) receives the number of busy slots.  With inuse NULL (what
 * nln/mac.c passes) slots are handed out round robin and polling stops at the
 * first idle one. */
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
/* nln/gravcuda.h:21 (impl nln/pgrav.cu:969-1020): runs the deferred calls and
 * waits for every slot, then btab[i*stride+7..10] += (ax, ay, az, phi)
 * accumulated since WalkInitSinkCUDA. */
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
/* Runs the deferred calls (below), then waits for the device. */
int pgrav_synchronize(void);
/* 1 (default): grav_mn_CUDA, pinteractCUDA, grav_mns_CUDA and grav_qns_CUDA
 * copy their lists into a pinned batch and return; a batch is uploaded in one
 * copy and evaluated by one kernel per kind when it holds 16 MB of lists, and
 * before WalkTerminateSinkCUDA (and every other call that reads the results or
 * replaces a table) returns.  Results are those of one launch per call, summed
 * in another order.  0: one copy and launch per call on the call's slot, as
 * nln/pgrav.cu:1043-1084 does.  A call of more than 65,536 tasks (64 sinks by
 * 128-512 sources each) or of more than 16 MB of list runs at once in either
 * mode.  Returns the previous setting. */
int pgrav_set_dropin_deferral(int enable);

/* ------------------------------------------------------------------ part 3
 * Batched path.  No single reference counterpart: it replaces the host walk's
 * stream of grav_mn_CUDA / grav_mnss_CUDA / grav_qnss_CUDA calls (nln/mac.c:
 * 352-475, 877-1272) for one force evaluation by one launch over every sink
 * group.  The tree, moments and lists are built on the host (as nln/cofm.c
 * and nln/mac.c do) inside pgrav_plan_create. */
typedef struct {
    double theta;          /* opening angle (Bmax MAC, nln/cofm.c:568-569) */
    int group_max;         /* bodies per sink group, <= 64 */
    int leaf_max;          /* bodies per tree leaf */
    int order;             /* cell multipole order: 0, 2 or 4 */
    int pcut;              /* cells with fewer bodies are evaluated body by body (0: auto) */
    int soft_type;         /* 0 none, 1 Plummer, 4 Dehnen K1 (nln/mac.c:196-237) */
    double eps;            /* softening length */
    double G;              /* Newton's constant applied to the results (nln/mac.c:269-277) */
    int periodic;          /* 0 isolated; 1: nearest image + Ewald correction table per pair
                              (nln/ewald.c); 2: the (2 nimage + 1)^3 images listed explicitly
                              (nln/mac.c:91-111) and the far images by the cubic local
                              expansion + background (nln/main.c:1811-1862, ewald_le2.c) */
    double box;            /* periodic box edge */
    int nthreads;          /* host threads for the build (0: all) */
    int merge_groups;      /* sink groups: 0 maximal cells; 1 (default) small siblings merged;
                              4 then Morton neighbours merged where cheaper; 5 balanced
                              coordinate bisection of cells into groups of ~group_max bodies
                              (groups are body sets, read through a slot table) */
    int persistent_blocks; /* 0: one workgroup per group; >0: persistent grid of that many
                              single-wave workgroups; -1 (default): 32 per CU */
    int fp64;              /* also build the FP64 tables (variants 3 and 4) */
    int dispatch;          /* 0: global longest-first; per-XCD Morton slices of equal estimated
                              cost, ordered inside 1: longest-first, 2 (default): Morton order,
                              3: the costliest 2 % first, then Morton order */
    int mixed_order;       /* 1: list an accepted cell at the lowest order whose truncation
                              bound (bmax/d)^(q+1) is within the order-`order` bound at the
                              opening-angle boundary, theta^(order+1): order 2 beyond
                              bmax * theta^(-(order+1)/3), order 0 beyond bmax * theta^-(order+1)
                              (the order-by-distance choice of nln/mac.c:1797-1810); 0: off */
    int nimage;            /* periodic = 2: explicit images per side, 0..2 (default 2,
                              nln/main.c:246) */
    int rcb_cell;          /* merge_groups = 5: tree cells of at most this many bodies are
                              cut into balanced groups by coordinate bisection (default 32768) */
    int mac;               /* acceptance test: 0 (default) opening angle theta; 1 error
                              tolerance, the reference's AREL_MAC: per cell, critical radii of
                              orders 0, 2, 4 from the Salmon-Warren bound (nln/cofm.c:276-565),
                              a cell taken at the lowest order whose radius + bmax the group's
                              nearest sink clears (DLRcritMAC, nln/mac.c:1776-1800); theta,
                              pcut and mixed_order are then unused */
    double errtol;         /* mac 1: absolute acceleration error per interaction, G = 1 units
                              (the reference's this_tol, nln/main.c:844-858); > 0 */
    double frac_tol;       /* mac 1: error relative to the cell's own m / r^2 (0: off) */
    double frac_tol0;      /* mac 1: relative floor: no radius beyond this one's (0: off) */
    double ptol_boost;     /* mac 1: errtol / (1 + ptol_boost bmax / r0), r0 the half-extent */
    double stol_max;       /* mac 1: > 1 scales errtol by the reference's sigma(m) table */
    int quad_ncut;         /* mac 1: cells of at least this many bodies may be taken at order 2
                              (default 7, nln/main.c:540) */
    int hexa_ncut;         /* mac 1: ... at order 4 (default 20, nln/main.c:541) */
    int body_unroll;       /* rel32c kernels: unroll of the order-0 and body loops, 2, 4 or 8
                              (0: 4) -- a performance choice, same results */
    double ewald_ratio;    /* periodic 1: a source's Ewald correction enters its group's
                              expansion (value, Jacobian and second derivatives at the
                              source, summed per group once and evaluated per sink, no
                              per-pair lookup) when the group radius is at most
                              ewald_ratio times the distance from the group to the
                              source's nearest other image and every sink sees the same
                              image; otherwise it is looked up per pair.  Default 0.05;
                              0: always per pair */
    int group_box;         /* 0 (default): a cell's distance to a group is measured to the
                              group's enclosing sphere; 1: to its bounding box (every sink
                              is inside it: a tighter bound for every sink, fewer cells
                              opened, larger errors at the same theta) */
} pgrav_params;

typedef struct {
    int64_t n, nnodes, ngroups;
    int64_t pairs[4];   /* (sink, source) pairs: order-4 cells, order-2 cells, order-0 cells, bodies */
    int64_t entries[4]; /* list entries per kind */
    double t_tree, t_lists, t_upload; /* host seconds */
    double mean_group;  /* bodies per group */
} pgrav_stats;

/* variants of pgrav_plan_evaluate */
#define PGRAV_REL32C 0  /* FP32, group-relative coordinates, compensated sums (default) */
#define PGRAV_REL32 1   /* FP32, group-relative, plain sums */
#define PGRAV_ABS32 2   /* FP32, absolute coordinates, plain sums */
#define PGRAV_F64 3     /* FP64, one batched launch */
#define PGRAV_PLAIN64 4 /* FP64, one launch per group on 32 streams */

void pgrav_params_default(pgrav_params *p);
/* pos3: n x 3 FP64 positions, mass: n FP64.  Returns NULL on error. */
void *pgrav_plan_create(const double *pos3, const double *mass, int64_t n, const pgrav_params *p);
void pgrav_plan_destroy(void *plan);
/* New positions and masses (caller order) with the tree topology, sink groups
 * and lists held fixed: moments and centres are recomputed in FP64 and the
 * device tables refreshed (what the adjoint's "tree held fixed" means).
 * Periodic plans keep each body at the image nearest its current position (a
 * body that crosses the box boundary stays beside its cell).  Waits for the
 * plan's last evaluation first; returns when the update is done.  Evaluations
 * of one plan must not overlap one another or an update. */
int pgrav_plan_update(void *plan, const double *pos3, const double *mass);
/* The same from device arrays (caller order, FP64), ready on `stream` (NULL: the
 * plan's stream). */
int pgrav_plan_update_device(void *plan, const double *pos3, const double *mass, void *stream);
/* Adjoint of the REL32C evaluation with the tree and lists held fixed: for
 * L = sum_i w_i . a_i + v_i phi_i (w4[i] = (w_i, v_i), caller order) returns
 * grad4[i] = (dL/dx_i, dL/dy_i, dL/dz_i, dL/dm_i).  Fixed-order sums: two calls
 * give identical gradients.  No reference counterpart (the reference has no
 * derivatives). */
int pgrav_plan_adjoint(void *plan, const double *w4, double *grad4);
/* The same on device arrays in caller order: w4 float (w_fp64 = 0) or double
 * quadruples, ready on `stream`; grad4 double quadruples, written in order on
 * `stream` (NULL: the plan's stream) when this returns.  Arrays aligned to one
 * quadruple. */
int pgrav_plan_adjoint_device(void *plan, const void *w4, int w_fp64, double *grad4, void *stream);
int pgrav_plan_stats(void *plan, pgrav_stats *st);
/* Launch shape for later evaluations: 0 one workgroup per group, -1 auto
 * persistent grid, k > 0 a persistent grid of k workgroups (a captured graph
 * keeps the shape it was captured with). */
int pgrav_plan_set_launch(void *plan, int persistent_blocks);
/* Unroll factor of the order-0 and body loops of the rel32c kernels for later
 * evaluations: 2, 4 (default) or 8 (an autotuner choice; same results; a captured
 * graph keeps the kernel it was captured with). */
int pgrav_plan_set_unroll(void *plan, int unroll);
/* List entries staged in LDS per round by the rel32c kernels for later
 * evaluations: 32, 64 (default) or 128 (an autotuner choice: one barrier pair
 * and one compensated fold per round, so results differ from 64's in the last
 * bits only).  Tiles of 32 and 128 run with the unroll of 4; periodic-1 plans
 * keep 64.  A captured graph keeps the kernel it was captured with. */
int pgrav_plan_set_tile(void *plan, int tile);
/* Group dispatch order (pgrav_params.dispatch, 0..3) for later evaluations. */
int pgrav_plan_set_dispatch(void *plan, int mode);
/* Enqueues one force evaluation on `stream` (NULL: the plan's stream). */
int pgrav_plan_evaluate(void *plan, int variant, void *stream);
/* Average device time of one evaluation over `reps` back-to-back launches,
 * bracketed by HIP events on the launch stream. */
int pgrav_plan_time(void *plan, int variant, int reps, void *stream, double *ms_per_launch);
int pgrav_plan_graph_capture(void *plan, int variant);
int pgrav_plan_graph_launch(void *plan, void *stream);
int pgrav_plan_synchronize(void *plan);
/* (ax, ay, az, phi) x n in the caller's body order, after G scaling. */
int pgrav_plan_result(void *plan, int variant, double *acc4);
/* The same into a device array (float quadruples, or double when fp64 = 1,
 * aligned to one quadruple), queued on `stream` (NULL: the plan's stream) after
 * the plan's last evaluation. */
int pgrav_plan_result_device(void *plan, int variant, void *acc4, int fp64, void *stream);
/* FP64 direct sum (nln/grav_n2.c pattern, same softening) for `ns` bodies
 * given by their original indices; out4 is ns x 4.  Periodic plans: nearest
 * image plus the Ewald correction interpolated from the FP64 table. */
int pgrav_plan_direct(void *plan, const int64_t *sample, int ns, double *out4);
/* Device pointers of the FP32 / FP64 results and the sorted FP64 bodies. */
int pgrav_plan_device_buffers(void *plan, void **out32, void **out64, void **bpos);
/* The device's FP64 centre of mass + mass, bmax and 35 raw moments of nodes
 * [node0, node0 + count) (after pgrav_plan_update: as recomputed on the GPU). */
int pgrav_plan_cells(void *plan, int64_t node0, int64_t count, double *cofm4, double *bmax,
                     double *raw35);
/* perm[i] = original index of sorted slot i. */
int pgrav_plan_permutation(void *plan, int64_t *perm);

/* Host-only tree and lists (no device): the exact lists a plan evaluates, for
 * inspection and for checking against the oracle.  List kinds: 0 order-4
 * cells, 1 order-2 cells, 2 order-0 cells (node indices), 3 bodies (sorted
 * slots). */
void *pgrav_host_plan(const double *pos3, const double *mass, int64_t n, const pgrav_params *p);
void pgrav_host_destroy(void *h);
int pgrav_host_stats(void *h, pgrav_stats *st);
int pgrav_host_group(void *h, int64_t g, int64_t *first, int *count, double *center4);
int64_t pgrav_host_list(void *h, int64_t g, int kind, uint32_t *out, int64_t cap);
/* periodic = 2: the image of each entry of that list, (i + n)(2n+1)^2 + (j + n)(2n+1)
 * + (k + n) for the source offset (i, j, k) x box (nln/mac.c:91-111 numbering). */
int64_t pgrav_host_list_images(void *h, int64_t g, int kind, uint8_t *out, int64_t cap);
/* The sorted body slots of group g (count of them returned; `slots` may be NULL):
 * the run first .. first + count, or for merge_groups = 5 the group's body set. */
int64_t pgrav_host_group_slots(void *h, int64_t g, int64_t *slots);
/* All groups at once: off[ngroups + 1] and slots[n], group g's sinks being
 * slots[off[g] .. off[g + 1]). */
int pgrav_host_group_slot_table(void *h, int64_t *off, int64_t *slots);
/* All groups at once: first body, count and FP64 centre + radius (arrays of ngroups). */
int pgrav_host_groups(void *h, int64_t *first, int *count, double *center4);
/* One list kind for all groups as CSR: off[ngroups+1], idx[off[ngroups]]; returns
 * the entry count (either pointer may be NULL). */
int64_t pgrav_host_csr(void *h, int kind, int64_t *off, uint32_t *idx);
int pgrav_host_cell(void *h, int64_t node, double *cofm3, double *bmax, double *raw35,
                    int64_t *first, int *count, int *child, int *nchild);
/* mac = 1 plans: the critical radii of orders 0, 2, 4 of every node (3 per node,
 * NaN where the node has too few bodies for that order); -1 for mac = 0 plans. */
int pgrav_host_rcrit(void *h, double *rcrit3);
/* The critical radii of one cell from its mass, extent bmax, absolute moments
 * B2, B4, B6 (sum m |x - c|^k about the centre of mass) and body count, under the
 * tolerances of p (mac, errtol, frac_tol, frac_tol0, ptol_boost, stol_max, quad_ncut,
 * hexa_ncut); r0: the system half-extent ptol_boost divides by.  nln/cofm.c:276-565. */
int pgrav_host_arel_rcrit(double m, double bmax, double B2, double B4, double B6, int64_t count,
                          double r0, const pgrav_params *p, double *rcrit3);
/* 29-field record of a node (nln/cofm.c:398-562 normalisation, R = bmax). */
int pgrav_host_record(void *h, int64_t node, double *rec29);
int pgrav_host_records(void *h, const uint32_t *nodes, int64_t n, double *rec29);
int pgrav_host_sorted(void *h, int64_t *perm, double *pos3);
/* Far field of the explicit-image scheme on the host, in FP64 (the check for
 * the device pass of periodic = 2 plans): for bodies in [0, box)^3, acc4[i] =
 * the images beyond the (2 nimage + 1)^3 listed ones (cubic local expansion
 * to degree 8) plus the uniform-background terms, times G -- what
 * FixCubeEwaldLE adds (nln/main.c:1836-1860; ewald_le, ewald_background in
 * nln/ewald_le2.c).  pgrav_host_le_lattice returns the lattice coefficient c_k
 * (k = 4, 6, 8): the x^k coefficient of the far images' 1/r sum. */
int pgrav_host_ewald_le(const double *pos3, const double *mass, int64_t n, double box, int nimage,
                        double G, double *acc4);
/* The same pass in the subtract_background form (FixCubeEwaldLE with
 * subtract_background, nln/main.c:1839-1860; ewald_le2.c:371-375, :605-646): the
 * far images of the bodies less the box's uniform background, the acceleration
 * less (4 pi / 3) G / L^2 times the dipole about the box centre; no background or
 * self terms.  For the lists of pgrav_dl_walk with subtract_background = 1. */
int pgrav_host_ewald_le_sb(const double *pos3, const double *mass, int64_t n, double box,
                           int nimage, double G, double *acc4);
/* Both forms on the GPU for caller device arrays (no reference counterpart: the
 * reference runs FixCubeEwaldLE on the host): pos3 (n x 3), mass (n) and acc4 (n x 4)
 * are FP64 device pointers, bodies in [0, box)^3; acc4 += the far field times G
 * (the same terms as pgrav_host_ewald_le, or pgrav_host_ewald_le_sb with
 * subtract_background = 1), queued on `stream` (a hipStream_t; NULL: the null
 * stream).  Scratch is stream-ordered (hipMallocAsync), the lattice matrix is
 * built once per device and form. */
int pgrav_ewald_le_device(const double *pos3, const double *mass, int64_t n, double box,
                          int nimage, double G, int subtract_background, double *acc4,
                          void *stream);
int pgrav_host_le_lattice(int k, int nimage, double *ck);
int pgrav_host_lpt(void *h, int32_t *order);

/* Keys of the reference's tree (host-only).  A key is 1 followed by 31 bits per
 * axis interleaved x y z from the top bit down (94 bits), returned as two words
 * key2 = (low 64 bits, high bits), the reference's two-word Key_t.
 * pgrav_host_fix_rsize: the key cube from the bounding box, nln/physics_generic.c:77-102
 * (FixRsize, exact = 0) or :104-128 (FixRsizeExact, exact = 1).
 * pgrav_host_keys: kind 0 GetKey (Morton, :146-167; bodies outside the cube are an
 * error, as IntPos :130-144 makes them), 1 GetKeyFast (:172-223, same keys for bodies
 * inside the cube, no check), 2 GetKeyPH (Peano-Hilbert, :226-233 over the automaton
 * of :340-482).  Positions are float, as in the reference's body.
 * pgrav_host_cell_corner: the corner of the cell a key names, CellCorner (:236-258,
 * size = the cell's edge) or, ph_curve = 1, CellCornerPH (:260-272, size = the finest
 * cell's edge, as there; the corner from the exact inverse of the PH key). */
int pgrav_host_fix_rsize(const float *rmin, const float *rmax, int exact, float *out_rmin,
                         float *rsize, float *keyfactor);
int pgrav_host_keys(const float *pos3, int64_t n, const float *rmin, float rsize, float keyfactor,
                    int kind, uint64_t *key2);
int pgrav_host_cell_corner(const uint64_t *key2, int ph_curve, const float *rmin, float rsize,
                           float *corner3, float *size);

/* The reference's list producer for the drop-in path (host-only): the DL walk
 * with DLRcritMAC (nln/mac.c:1773-1844) over a tree of one body per leaf, the
 * sink inheritance of InheritSinkNlogN (nln/mac.c:477-872) and the M x N calls
 * of mxn_hexa / mxn_quad2 (nln/mac.c:941-1272) with the pQ1 queue
 * (nln/mac.c:394-475).  Emits what the reference hands nln/gravcuda.h: SoA-8
 * blocks of absolute-position records (Hvec 29 floats, Qvec 13, Mvec 4 per
 * record, nln/mac.c:117-153) and the calls that consume them, plus each body's
 * remainder lists (the reference's CPU leaf, nln/mac.c:506-700).  The walk
 * that drives the MAC (libtree's WalkNT) is not in the reference checkout:
 * this one is depth first in Morton order, one MAC call per sink with every
 * pending source.  pgrav_params: order, mac (0 opening angle, 1 error
 * tolerance radii), quad_ncut / hexa_ncut (p2cut / p4cut), periodic 0 or 2
 * (explicit images, nimage, box); leaf_max is 1 (the reference's tree). */
typedef struct {
    float dlfac;      /* DLfac (default 6, nln/main.c:536) */
    int leaf_max_n;   /* default 56 = quad_ncut * 8 (nln/main.c:544) */
    int gated;        /* 0: the reference's terminal expansion of every SPLIT_SRC cell into
                         its bodies (nln/mac.c:1818-1832); 1 (default): only cells of at
                         most leaf_max_n bodies, larger ones opened by the walk (the gate of
                         DLRcritMACsb, nln/mac.c:1740) */
    int hblock, min_qsink, min_hsink, min_qsrc, min_hsrc; /* mxn_s, defaults mxn_cuda */
    int do_pQ, do_pQL, do_pH;                              /* (nln/main.c:215-225) */
    int log_calls;    /* 1: record every MAC call (sink, sources, results, appends) */
    /* subtract_background (nln/main.c:543): 1 selects DLRcritMACsb (nln/mac.c:1565-1771)
       with periodic 2: cell records carry the moments less the cell's uniform cube of the
       mean density (mpole_add_cube, nln/cofm.c:348, :626-648), the walk adds anti-masses
       for the background of split cells, empty octants and bodies outside each sink's
       near cube, the leaf adds the near cube's own field (cubic_acc, nln/mac.c:688-694),
       and small central-image cells become MMvec segments (nln/mac.c:1736-1741) batched by
       mxn_mono into the pMM_mnss_sK1 queue (nln/mac.c:352-457, :877-936). */
    int subtract_background;
    int do_pM, min_msink, min_msrc;                        /* mxn_cuda: 1, 8, 128 */
    float nx;         /* mac.nx, the cube root of the run's global body count (nln/main.c:856),
                         which sets where near cubes start (nln/mac.c:747); 0: cbrt(n) */
    int geometric_center; /* geometric_center (nln/main.c:542): cells expanded about their
                         key cube's centre (dipole moments kept, nln/cofm.c:141-146) instead
                         of their centre of mass; the near cubes then align with the cells */
    int tail_spans;   /* 0: each body's leaf remainders copied whole into SOURCES (TAILS);
                         1: every sink's new stack blocks emitted once, each body's
                         remainders given as runs of them (TAIL_RUNS / TAIL_INDEX; TAILS then
                         holds -1 offsets and the record counts) -- memory O(appends)
                         instead of O(N x list length) */
} pgrav_dl_params;

/* arrays of pgrav_dl_array (element type, values per element) */
#define PGRAV_DL_CELLS 0        /* float 8 per node: mass, x, y, z, rcrit, bmax, rcrit_q, rcrit_h */
#define PGRAV_DL_RECORDS 1      /* float 29 per node: the record layout of Hvec (R = bmax) */
#define PGRAV_DL_BODIES 2       /* float 4 per sorted slot: mass, x, y, z */
#define PGRAV_DL_PERM 3         /* int64 1 per sorted slot: the caller's index */
#define PGRAV_DL_CHILD_OFF 4    /* int64 nnodes + 1: node c's daughters are children[off[c]..] */
#define PGRAV_DL_CHILDREN 5     /* int32: a node, or a body -(slot + 1) */
#define PGRAV_DL_CALLS 6        /* int64 6 per M x N call, in walk order: routine (0 pH,
                                   1 pQ), first body slot, m, source offset (floats into
                                   SOURCES), source_n, record size */
#define PGRAV_DL_SOURCES 7      /* float: every emitted SoA-8 block */
#define PGRAV_DL_QFLUSH 8       /* int64 6 per grav_qnss_CUDA("pQ1_qnss") flush: sink base,
                                   m, first segment (QSEG), source offset, source length, 0 */
#define PGRAV_DL_QSEG 9         /* uint32 2 per queued sink: base, length (records) */
#define PGRAV_DL_TAILS 10       /* int64 6 per sorted slot: the leaf's Hvec, Qvec, Mvec
                                   remainders as (source offset, records) pairs, padded to 8 */
#define PGRAV_DL_INTERACTIONS 11 /* int64 per sorted slot: bodies accounted for (== nimage N) */
#define PGRAV_DL_LOG_CALLS 12   /* int64 4 per MAC call: sink, first source, n, first append */
#define PGRAV_DL_LOG_SOURCES 13 /* int32 3 per source: entity, image, result (0 accept,
                                   1 split sink, 2 split source) */
#define PGRAV_DL_LOG_APPENDS 14 /* int32 3 per append: list (0 M, 1 Q, 2 H), entity, image */
#define PGRAV_DL_NODE_INFO 15   /* int64 4 per node: first body, bodies, has child nodes, level */
/* subtract_background = 1 only (empty otherwise) */
#define PGRAV_DL_NODE_GEOM 16   /* float 5 per node: key-cube corner x, y, z, edge, octants
                                   holding bodies (bit j: octant x = j & 1, y = j >> 1 & 1,
                                   z = j >> 2, the key digit of CellCorner) */
#define PGRAV_DL_UBS 17         /* float 4 per node: the unsubtracted mass, qxxxx, qyyyy,
                                   qxxyy (hexacell umass / uxxxx / uyyyy / uxxyy) */
#define PGRAV_DL_BODY_LEAF 18   /* int64 per sorted slot: the leaf node holding the body */
#define PGRAV_DL_MFLUSH 19      /* int64 8 per grav_mnss_CUDA("pMM_mnss_sK1") flush: sink
                                   base, m, first ss_index (MSS_INDEX), first ss entry
                                   (MSS_SEG / MSS_SOURCE_N), ss_len, first segment (MSEG),
                                   seg_n, 0 */
#define PGRAV_DL_MSS_INDEX 20   /* uint16 per queued sink */
#define PGRAV_DL_MSS_SEG 21     /* uint32 2 per ss entry: first segment, segments */
#define PGRAV_DL_MSS_SOURCE_N 22 /* int32 per ss entry: bodies in its segments */
#define PGRAV_DL_MSEG 23        /* uint32 2 per MMvec segment: first body slot, bodies */
#define PGRAV_DL_MCALLS 24      /* int64 6 per mxn_mono run on the host (not queued): first
                                   slot, m, first segment (MSEG), segments, bodies, 0 */
#define PGRAV_DL_MMTAILS 25     /* int64 3 per sorted slot: the leaf's MMvec remainder:
                                   first segment (MSEG), segments, bodies */
#define PGRAV_DL_BACKGROUND 26  /* double 3 per sorted slot: -rho0 cubic_acc(pos - cen, cr),
                                   the near cube's field, FP64 */
#define PGRAV_DL_SB_SINK 27     /* float 6 per sorted slot: the leaf sink's cen x, y, z, cr,
                                   fmass (signed mass appended to its lists), near */
#define PGRAV_DL_LOG_SB 28      /* float 8 per MAC call: the sink's cen x, y, z, cr, cr2,
                                   clevel (42: none yet), done, near after the call */
#define PGRAV_DL_LOG_FLAGS 29   /* int32 2 per logged source: background flag in, out */
#define PGRAV_DL_SB_STATS 30    /* int64 8: pMM_mnss_sK1 flushes, host mxn_mono runs, MMvec
                                   segments, body / cell / empty-octant anti-masses, queued
                                   sink-body pairs, 0 */
/* tail_spans = 1 only */
#define PGRAV_DL_TAIL_RUNS 31   /* int64 3 per run: list (0 Hvec, 1 Qvec, 2 Mvec), source
                                   offset (floats into SOURCES), records (SoA-8 blocks x 8) */
#define PGRAV_DL_TAIL_INDEX 32  /* int64 2 per sorted slot: first run, runs (H, then Q, then
                                   M, each in stack order) */
/* Appends in LOG_APPENDS with subtract_background: 1 / 2 Q / H with subtracted moments,
 * 3 / 4 Q / H of a background-flagged cell (Q mass + the cube's, H the unsubtracted
 * umass / uxxxx / uyyyy / uxxyy), 5 a body's anti-mass (-rho0 edge^3 of its leaf cube),
 * 6 a split cell's anti-mass (-ucell mass of its level), 8 + j the anti-mass of empty
 * octant j, 20 an MMvec segment (entity: the cell); every anti-mass sits at the source's
 * own position plus offset (appendMvec writes r, nln/mac.c:1421-1430). */

void pgrav_dl_params_default(pgrav_dl_params *d);
/* Runs the walk; NULL on error (pgrav_last_error). */
void *pgrav_dl_walk(const double *pos3, const double *mass, int64_t n, const pgrav_params *p,
                    const pgrav_dl_params *d);
void pgrav_dl_destroy(void *h);
/* out16: MAC calls, pH calls, pQ calls, pQ1 flushes, source floats, M x N calls, leaf
 * H / Q / M records, M x N H pairs, M x N Q pairs, bodies with a wrong interaction
 * count, n, nodes, images, logged calls. */
int pgrav_dl_stats(void *h, int64_t *out16);
/* Copies array `which` into out (NULL: size only); returns its element count. */
int64_t pgrav_dl_array(void *h, int which, void *out);

/* ------------------------------------------------------------------ part 4
 * The step either side of the force evaluation: the leapfrog kick-drift of the
 * body table.  Fields are float arrays with a common byte stride (the
 * reference passes &btab[0].mass, &btab[0].pos[0], ... and sizeof(body)); in
 * and out may be the same table.  Arithmetic is the reference's: FP64 steps,
 * results rounded to float, no fused multiply-adds. */
/* nln/integrate.h:5-20 (Integrate), same arguments, host pointers: the tables'
 * spans are copied to the GPU and back.  vel += dtv acc; pos += dt vel;
 * ke = 1/2 sum m |vel + (tpos - tvel) acc|^2, pe = 1/2 sum m phi; tvel += dtv,
 * tpos += dt. */
int pgrav_integrate(const float *inmass, const float *inpos, const float *invel, const float *inacc,
                    const float *inphi, int instride, float *outpos, float *outvel, int outstride,
                    int n, double dt, double dtv, double *tpos, double *tvel, double *kep,
                    double *pep);
/* The same on device pointers (a table resident in HBM), on `stream` (NULL: the
 * null stream); returns after kep / pep are known. */
int pgrav_integrate_device(const float *inmass, const float *inpos, const float *invel,
                           const float *inacc, const float *inphi, int instride, float *outpos,
                           float *outvel, int outstride, int n, double dt, double dtv, double *tpos,
                           double *tvel, double *kep, double *pep, void *stream);
/* No counterpart in the reference.  pgrav_integrate_device for callers that
 * cannot allocate (code under HIP graph capture, tight stepping loops, harnesses
 * that forbid hipMalloc): the same arguments and results, but no hipMalloc or
 * hipFree -- the per-workgroup energy partials and their total live in a static
 * device array of the library (8 MiB per GPU, held until the process ends), so
 * n is limited to 134,217,728 bodies (larger n is an error).  Calls that overlap
 * on one GPU share that array and race: serialise them. */
int pgrav_integrate_device_noalloc(const float *inmass, const float *inpos, const float *invel,
                                   const float *inacc, const float *inphi, int instride,
                                   float *outpos, float *outvel, int outstride, int n, double dt,
                                   double dtv, double *tpos, double *tvel, double *kep, double *pep,
                                   void *stream);
/* nln/integrate.c:82-135 (CosmoIntegrate) on device pointers, with the
 * cosmology's factors passed in (the reference computes them from its
 * cosmology object, libcosmo, which is not part of this path):
 * a0 = a(tpos), a1 = a(tpos + dt), dt_kick_topos = a0^2 kick(tvel, tpos),
 * dt_kick = a0^2 kick(tvel, tvel + dtv), dt_drift = drift(tpos, tpos + dt).
 * The caller advances tpos and tvel. */
int pgrav_cosmo_integrate_device(const float *inmass, const float *inpos, const float *invel,
                                 const float *inacc, const float *inphi, int instride,
                                 float *outpos, float *outvel, int outstride, int n, double a0,
                                 double a1, double dt_kick_topos, double dt_kick, double dt_drift,
                                 double *kep, double *pep, void *stream);
/* nln/integrate.c:62-80 (ConvertV) on a device table: vel /= scale
 * (to_physical) or vel *= scale. */
int pgrav_convert_v_device(float *vel, int stride, int n, double scale, int to_physical,
                           void *stream);

/* ------------------------------------------------------------------ part 5
 * SDF snapshots of the body table (the format nln/output.c:270-423 writes
 * through libSDF's SDFwrite64, laid out as sph+nln/wind.c:82-135 writes it):
 * an ASCII header of typed declarations and a struct description, then the
 * packed records.  Host-only. */
#define PGRAV_SDF_INT 0
#define PGRAV_SDF_INT64 1
#define PGRAV_SDF_FLOAT 2
#define PGRAV_SDF_DOUBLE 3
#define PGRAV_SDF_STRING 4
#define PGRAV_SDF_CHAR 5  /* struct members only */
#define PGRAV_SDF_SHORT 6 /* struct members only */
typedef struct {
    const char *name;
    int type;        /* PGRAV_SDF_INT .. PGRAV_SDF_STRING */
    int64_t i;       /* INT, INT64 */
    double d;        /* FLOAT, DOUBLE */
    const char *s;   /* STRING */
} pgrav_sdf_value;
/* Writes the header values in order, `desc` ("struct {...}") with [nrec], and
 * nrec records of recsize bytes. */
int pgrav_sdf_write(const char *path, const pgrav_sdf_value *vals, int nvals, const char *desc,
                    const void *records, int64_t nrec, int64_t recsize);
/* Parses the header; NULL on error.  Strings returned stay valid until close. */
void *pgrav_sdf_open(const char *path);
void pgrav_sdf_close(void *h);
int pgrav_sdf_nvalues(void *h);
int pgrav_sdf_value_at(void *h, int k, pgrav_sdf_value *out);
int64_t pgrav_sdf_nrecords(void *h);
int64_t pgrav_sdf_recsize(void *h);
int pgrav_sdf_nfields(void *h);
/* Member k of the record struct (arrays are expanded to name[i]): byte offset, type. */
int pgrav_sdf_field_at(void *h, int k, const char **name, int *type, int64_t *offset);
int pgrav_sdf_read(void *h, int64_t first, int64_t count, void *out);

#ifdef __cplusplus
}
#endif
#endif
