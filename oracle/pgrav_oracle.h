/*
 * TEST INFRASTRUCTURE ONLY.  FP64 CPU restatement of the 2HOT gravity
 * interaction-list evaluation, used solely as the parity checker by tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py.  Nothing in the
 * product library links, loads or calls this code.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the lanl/ahotsph-2hot-apps checkout).
 */
#ifndef PGRAV_ORACLE_H
#define PGRAV_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Field order of one source record (nln/pgrav.cu:16-47). */
enum {
    ORC_MASS = 0, ORC_X, ORC_Y, ORC_Z, ORC_R,
    ORC_QX, ORC_QY, ORC_QZ,
    ORC_QXX, ORC_QXY, ORC_QYY, ORC_QXZ, ORC_QYZ,
    ORC_QXXX, ORC_QXXY, ORC_QXYY, ORC_QYYY, ORC_QXXZ, ORC_QXYZ, ORC_QYYZ,
    ORC_QXXXX, ORC_QXXXY, ORC_QXXYY, ORC_QXYYY, ORC_QYYYY, ORC_QXXXZ, ORC_QXXYZ, ORC_QXYYZ,
    ORC_QYYYZ,
    ORC_HSZ
};

/* One multipole source evaluated in FP64 at displacement (x,y,z) = sink - source.
 * order 0/2/4; dipole on/off; eps2 added to r^2 (0 for the GPU kernels);
 * quad_phi_sign = +1 (physical, do_gravph/pgrav_vec.c) or -1 (pgrav.cu pH/pQ,
 * do_gravpq: "phi -= eqe").  Accumulates (ax, ay, az, phi) into acc. */
void orc_mpole(const double *q, int order, int dipole, double x, double y, double z, double eps2,
               int quad_phi_sign, double acc[4]);

/* m FP64 sinks against nc AoS 29-double records (sum of orc_mpole). */
void orc_eval_records(const double *sink_pos3, int64_t m, const double *rec29, int64_t nc,
                      int order, int dipole, double eps2, int quad_phi_sign, double *out4);

/* pM / pQ / pH (nln/pgrav.cu:49-410) and the CPU twins do_gravp/pq/ph
 * (nln/do_grav_sse4.c:1768-2224) over an SoA-nsse source list. */
void orc_pmn(int order, int dipole, const float *sink_pos3, int m, const float *f, int source_n,
             int sz, int nsse, double eps2, int quad_phi_sign, double *out4);

/* pM_sK1 (nln/pgrav.cu:412-468): monopoles with Dehnen K1 softening. */
void orc_pm_sk1(const float *sink_pos3, int m, const float *f, int source_n, int nsse,
                double eps_inv, double *out4);

/* pMM_sK1 (nln/pgrav.cu:575-637): every sink against the bodies named by a
 * segment list {base,length} into src_pos3 (float3 per body). */
void orc_pmm_sk1(const float *sink_pos3, int m, const uint32_t *seg, int seg_n,
                 const float *src_pos3, double mmass, double eps_inv, double *out4);

/* pMM1_mnss_sK1 (nln/pgrav.cu:713-790): sink i uses segment run ss_seg[ss_index[i]].
 * Returns the number of sinks whose summed segment length != source_n[...]. */
int orc_pmm1_mnss_sk1(const float *sink_pos3, int m, const uint16_t *ss_index,
                      const uint32_t *ss_seg, int ss_len, const uint32_t *seg, int seg_len,
                      const float *src_pos3, const int *source_n, double mmass, double eps_inv,
                      double *out4);

/* pQ1_mnss (nln/pgrav.cu:293-368): sink i uses its own run seg[i] of a packed
 * SoA-8 order-2 source buffer. */
void orc_pq1_mnss(const float *sink_pos3, int m, const uint32_t *seg, const float *ff,
                  int quad_phi_sign, double *out4);

/* pQQ1 (nln/pgrav.cu:498-572), with the reference's index defects fixed:
 * every sink against cells idx[0..source_n) of an AoS table of 13 floats. */
void orc_pqq1(const float *sink_pos3, int m, const int *idx, int source_n, const float *qtab13,
              int quad_phi_sign, double *out4);

/* FP64 direct sum (pattern of nln/grav_n2.c:162-173, non-periodic):
 * soft_type 0 = none, 1 = Plummer (r^2+eps^2), 4 = Dehnen K1 with radius eps
 * (nln/pgrav.cu:760-766).  The self pair (r == 0) is skipped when unsoftened
 * and kept (finite self potential) when softened. */
void orc_direct(const double *sink_pos3, int64_t m, const double *src_pos3,
                const double *src_mass, int64_t n, int soft_type, double eps, double *out4,
                int nthreads);

/* FP64 gradient of L = sum_i w_i.a_i + v_i phi_i of the direct sum (soft_type 0
 * or 1) with respect to every position and mass; w4[i] = (w_i, v_i). */
void orc_direct_grad(const double *pos3, const double *mass, int64_t n, const double *w4,
                     int soft_type, double eps, double *grad4, int nthreads);

/* Classic Ewald sum (nln/ewald.c:13-64): x = source - sink, box L; returns the
 * force per unit mass on the sink (f) and the potential (phi). */
void orc_ewald(const double x[3], double L, double f[3], double *phi);

/* Periodic FP64 direct sum with ewald() per pair (nln/grav_n2.c:162-173). */
void orc_direct_ewald(const double *sink_pos3, int64_t m, const double *src_pos3,
                      const double *src_mass, int64_t n, double L, double *out4, int nthreads);

/* The Ewald correction of n pairs and its derivatives.  r = source - sink (the
 * nearest image); C(r) = f(r) - r/|r|^3 and psi(r) = phi(r) + 1/|r| where f, phi
 * are the lattice sums of nln/ewald.c:13-64 (orc_ewald), i.e. the pull of a unit
 * source's periodic images and background with the nearest Newtonian term taken
 * out.  val4 = (C, psi); jac12 = dC_i/dr_j (rows i = 0..2) then dpsi/dr_j (= C_j);
 * third10 (may be NULL) = d2C_i/dr_j dr_k, fully symmetric, components xxx, xxy,
 * xxz, xyy, xyz, xzz, yyy, yyz, yzz, zzz; all differentiated term by term in
 * closed form.
 *   nodes == 0: at r.
 *   nodes  > 0: tabulated on the grid of spacing h = (L/2)/(nodes - 1): the node
 *               nearest r per axis (signed) and its Taylor steps to r = node + d:
 *               psi + C.d + d.J.d/2 + T d d d/6, C + J d + T d d/2, J + T d, and
 *               the node's T. */
void orc_ewald_corr(const double *r3, int64_t n, double L, int nodes, double *val4,
                    double *jac12, double *third10, int nthreads);

/* Raw moments (FP64) of a point set about `center` to order 4, summed
 * directly (what nln/cofm.c:650-768 mpole_add_mono accumulates), followed by
 * the detraced, bmax-normalised record of nln/cofm.c:398-562.  rec receives
 * the 29-float source record with pos = center and R = bmax. */
void orc_cell_record(const double *pos3, const double *mass, int64_t n, const double center[3],
                     double bmax, double rec[ORC_HSZ]);

/* Raw moment tensor components (35 values: m, x,y,z, 6 second, 10 third,
 * 15 fourth order, in the order used by nln/cofm.c:660-720). */
void orc_raw_moments(const double *pos3, const double *mass, int64_t n, const double center[3],
                     double raw[35]);

/* nln/integrate.c:16-59 (Integrate): leapfrog kick-drift of strided float
 * fields, FP64 steps rounded to float; returns ke, pe and advances tpos, tvel. */
void orc_integrate(const float *inmass, const float *inpos, const float *invel, const float *inacc,
                   const float *inphi, int instride, float *outpos, float *outvel, int outstride,
                   int n, double dt, double dtv, double *tpos, double *tvel, double *kep,
                   double *pep);
/* nln/integrate.c:82-135 (CosmoIntegrate) with the cosmology's factors given. */
void orc_cosmo_integrate(const float *inmass, const float *inpos, const float *invel,
                         const float *inacc, const float *inphi, int instride, float *outpos,
                         float *outvel, int outstride, int n, double a0, double a1,
                         double dt_kick_topos, double dt_kick, double dt_drift, double *kep,
                         double *pep);
/* nln/integrate.c:62-80 (ConvertV). */
void orc_convert_v(float *vel, int stride, int n, double scale, int to_physical);

#ifdef __cplusplus
}
#endif
#endif
