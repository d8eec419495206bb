/*
 * TEST INFRASTRUCTURE ONLY -- FP64 restatement of the reference's
 * interaction arithmetic.  Used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker; never linked into the product.
 *
 * Pins (see tests/test_oracle.py):
 *   orc_mpole / orc_pmn      -- closed-form multipole convergence against the
 *                               FP64 direct sum of the cell's own bodies, the
 *                               single-body special case, the reference's own
 *                               compiled CPU twins (oracle/_ref) and golden
 *                               vectors made from them (tests/golden/).
 *   orc_pm_sk1 / orc_pmm*    -- brute force + K1 continuity at r = eps, golden
 *                               vectors from the reference's _pgravm_sK1.
 *   orc_direct               -- Newton's third law, two-body closed forms.
 *   orc_ewald                -- lattice constant 2.8372975 and symmetry
 *                               (nln/ewald_le2.c:38-43).
 *   orc_ewald_corr           -- == orc_ewald minus the Newtonian term; the
 *                               closed-form Jacobian == central differences;
 *                               psi(0) == 2 x 1.4186487533, dC/dr(0) ==
 *                               -4 pi / 3 (nln/ewald_le2.c:35-43); tabulated
 *                               form exact at the nodes, error O(h^2).
 *   orc_cell_record          -- parallel-axis identity / tracelessness.
 */
#include "pgrav_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* nln/pgrav.cu:66 -- field k of source i in an SoA-nsse block layout */
static inline double fld(const float *f, int i, int k, int sz, int nsse) {
    return (double)f[(i / nsse) * nsse * sz + k * nsse + i % nsse];
}

/* nln/pgrav.cu:66-203 (pH body; pQ stops after the quadrupole, pM after the
 * monopole) and nln/do_grav_sse4.c:1798-1889 (do_gravph: softened, no dipole). */
void orc_mpole(const double *q, int order, int dipole, double x, double y, double z, double eps2,
               int quad_phi_sign, double acc[4]) {
    double r2 = x * x + y * y + z * z + eps2;
    double rinv = -1.0 / sqrt(r2); /* -rsqrtf(r2), pgrav.cu:72 */
    double t = rinv;
    double eqe = q[ORC_MASS] * t;
    double rinv2 = t * t;
    double eq0, eq1, eq2;
    acc[3] += eqe;
    eqe *= rinv2;
    acc[0] += x * eqe;
    acc[1] += y * eqe;
    acc[2] += z * eqe;
    if (order == 0)
        return;
    const double R = q[ORC_R];

    if (dipole) { /* pgrav.cu:83-94 */
        t *= R * rinv2;
        eq0 = q[ORC_QX] * t;
        eq1 = q[ORC_QY] * t;
        eq2 = q[ORC_QZ] * t;
        eqe = eq0 * x + eq1 * y + eq2 * z;
        acc[3] += eqe;
        eqe *= 3.0 * rinv2;
        acc[0] += x * eqe - eq0;
        acc[1] += y * eqe - eq1;
        acc[2] += z * eqe - eq2;
        t *= 3.0 * R * rinv2;
    } else { /* do_grav_sse4.c:1836-1837 */
        t *= rinv2;
        t *= 3.0 * rinv2 * R * R;
    }

    /* quadrupole, pgrav.cu:96-120 */
    eq0 = q[ORC_QXX] * x + q[ORC_QXY] * y + q[ORC_QXZ] * z;
    eq1 = q[ORC_QYY] * y + q[ORC_QXY] * x + q[ORC_QYZ] * z;
    eq2 = -(q[ORC_QXX] + q[ORC_QYY]) * z + q[ORC_QXZ] * x + q[ORC_QYZ] * y;
    eq0 *= t;
    eq1 *= t;
    eq2 *= t;
    eqe = 0.5 * (eq0 * x + eq1 * y + eq2 * z);
    acc[3] += quad_phi_sign * eqe;
    eqe *= 5.0 * rinv2;
    acc[0] += x * eqe - eq0;
    acc[1] += y * eqe - eq1;
    acc[2] += z * eqe - eq2;
    if (order == 2)
        return;

    /* octupole, pgrav.cu:122-161 */
    t *= 5.0 * R * rinv2;
    double xx = 0.5 * x * x, xy = x * y, xz = x * z;
    double yy = 0.5 * y * y, yz = y * z, zz = 0.5 * z * z;
    double xxx = ((1.0 / 3.0) * xx - zz) * x;
    double xxz = (xx - (1.0 / 3.0) * zz) * z;
    double yyy = ((1.0 / 3.0) * yy - zz) * y;
    double yyz = (yy - (1.0 / 3.0) * zz) * z;
    xx -= zz;
    yy -= zz;
    eq0 = q[ORC_QXXX] * xx + q[ORC_QXYY] * yy + q[ORC_QXXY] * xy + q[ORC_QXXZ] * xz
          + q[ORC_QXYZ] * yz;
    eq1 = q[ORC_QXYY] * xy + q[ORC_QXXY] * xx + q[ORC_QYYY] * yy + q[ORC_QYYZ] * yz
          + q[ORC_QXYZ] * xz;
    eq2 = -(q[ORC_QXXX] + q[ORC_QXYY]) * xz - (q[ORC_QXXY] + q[ORC_QYYY]) * yz
          + q[ORC_QXXZ] * xx + q[ORC_QYYZ] * yy + q[ORC_QXYZ] * xy;
    eq0 *= t;
    eq1 *= t;
    eq2 *= t;
    eqe = (1.0 / 3.0) * (eq0 * x + eq1 * y + eq2 * z);
    acc[3] += eqe;
    eqe *= 7.0 * rinv2;
    acc[0] += x * eqe - eq0;
    acc[1] += y * eqe - eq1;
    acc[2] += z * eqe - eq2;

    /* hexadecapole, pgrav.cu:163-203 */
    t *= 7.0 * R * rinv2;
    double xxy = y * xx, xyy = x * yy, xyz = xy * z;
    eq0 = q[ORC_QXXXX] * xxx + q[ORC_QXYYY] * yyy + q[ORC_QXXXY] * xxy + q[ORC_QXXXZ] * xxz
          + q[ORC_QXXYY] * xyy + q[ORC_QXXYZ] * xyz + q[ORC_QXYYZ] * yyz;
    eq1 = q[ORC_QXYYY] * xyy + q[ORC_QXXXY] * xxx + q[ORC_QYYYY] * yyy + q[ORC_QYYYZ] * yyz
          + q[ORC_QXXYY] * xxy + q[ORC_QXXYZ] * xxz + q[ORC_QXYYZ] * xyz;
    eq2 = -q[ORC_QXXXX] * xxz - (q[ORC_QXYYY] + q[ORC_QXXXY]) * xyz - q[ORC_QYYYY] * yyz
          + q[ORC_QXXXZ] * xxx + q[ORC_QYYYZ] * yyy - q[ORC_QXXYY] * (xxz + yyz)
          + q[ORC_QXXYZ] * xxy + q[ORC_QXYYZ] * xyy;
    eq0 *= t;
    eq1 *= t;
    eq2 *= t;
    eqe = 0.25 * (eq0 * x + eq1 * y + eq2 * z);
    acc[3] += eqe;
    eqe *= 9.0 * rinv2;
    acc[0] += x * eqe - eq0;
    acc[1] += y * eqe - eq1;
    acc[2] += z * eqe - eq2;
}

/* nln/pgrav.cu:49-410 (pH/pQ/pM) and nln/do_grav_sse4.c:1768-2224 */
void orc_pmn(int order, int dipole, const float *sink_pos3, int m, const float *f, int source_n,
             int sz, int nsse, double eps2, int quad_phi_sign, double *out4) {
    double q[ORC_HSZ];
    int nf = order == 0 ? 4 : (order == 2 ? 13 : 29);
    for (int s = 0; s < m; s++) {
        double acc[4] = {0, 0, 0, 0};
        for (int i = 0; i < source_n; i++) {
            memset(q, 0, sizeof(q));
            for (int k = 0; k < nf; k++) q[k] = fld(f, i, k, sz, nsse);
            orc_mpole(q, order, dipole, (double)sink_pos3[3 * s] - q[ORC_X],
                      (double)sink_pos3[3 * s + 1] - q[ORC_Y],
                      (double)sink_pos3[3 * s + 2] - q[ORC_Z], eps2, quad_phi_sign, acc);
        }
        for (int k = 0; k < 4; k++) out4[4 * s + k] = acc[k];
    }
}

/* m sinks (FP64) against nc AoS records of 29 doubles: the FP64 sum of
 * orc_mpole over a cell list. */
void orc_eval_records(const double *sink_pos3, int64_t m, const double *rec29, int64_t nc,
                      int order, int dipole, double eps2, int quad_phi_sign, double *out4) {
    for (int64_t s = 0; s < m; s++) {
        double acc[4] = {0, 0, 0, 0};
        for (int64_t c = 0; c < nc; c++) {
            const double *q = rec29 + 29 * c;
            orc_mpole(q, order, dipole, sink_pos3[3 * s] - q[ORC_X], sink_pos3[3 * s + 1] - q[ORC_Y],
                      sink_pos3[3 * s + 2] - q[ORC_Z], eps2, quad_phi_sign, acc);
        }
        for (int k = 0; k < 4; k++) out4[4 * s + k] = acc[k];
    }
}

/* Dehnen K1 pair term (nln/pgrav.cu:745-770, nln/py_pgrav/_pgrav.c:21-42) */
static void k1_pair(double x, double y, double z, double mass, double eps_inv, double acc[4]) {
    double eps_inv2 = eps_inv * eps_inv;
    double eps2 = 1.0 / eps_inv2;
    double r2 = x * x + y * y + z * z;
    double eqe;
    if (r2 > eps2) {
        double rinv = -1.0 / sqrt(r2);
        eqe = mass * rinv;
        acc[3] += eqe;
        eqe *= rinv * rinv;
    } else {
        double u2 = r2 * eps_inv2 - 1.0;
        double t = (((45.0 / 32.0) * u2 + (-3.0 / 8.0)) * u2 + (1.0 / 2.0)) * u2 - 1.0;
        t *= mass * eps_inv;
        acc[3] += t;
        eqe = ((-135.0 / 16.0) * u2 + (3.0 / 2.0)) * u2 - 1.0;
        eqe *= mass * eps_inv * eps_inv2;
    }
    acc[0] += x * eqe;
    acc[1] += y * eqe;
    acc[2] += z * eqe;
}

/* nln/pgrav.cu:412-468 */
void orc_pm_sk1(const float *sink_pos3, int m, const float *f, int source_n, int nsse,
                double eps_inv, double *out4) {
    for (int s = 0; s < m; s++) {
        double acc[4] = {0, 0, 0, 0};
        for (int i = 0; i < source_n; i++) {
            k1_pair((double)sink_pos3[3 * s] - fld(f, i, 1, 4, nsse),
                    (double)sink_pos3[3 * s + 1] - fld(f, i, 2, 4, nsse),
                    (double)sink_pos3[3 * s + 2] - fld(f, i, 3, 4, nsse), fld(f, i, 0, 4, nsse),
                    eps_inv, acc);
        }
        for (int k = 0; k < 4; k++) out4[4 * s + k] = acc[k];
    }
}

/* nln/pgrav.cu:575-637 */
void orc_pmm_sk1(const float *sink_pos3, int m, const uint32_t *seg, int seg_n,
                 const float *src_pos3, double mmass, double eps_inv, double *out4) {
    for (int s = 0; s < m; s++) {
        double acc[4] = {0, 0, 0, 0};
        for (int g = 0; g < seg_n; g++) {
            for (uint32_t j = 0; j < seg[2 * g + 1]; j++) {
                const float *qp = src_pos3 + 3 * (size_t)(seg[2 * g] + j);
                k1_pair((double)sink_pos3[3 * s] - qp[0], (double)sink_pos3[3 * s + 1] - qp[1],
                        (double)sink_pos3[3 * s + 2] - qp[2], mmass, eps_inv, acc);
            }
        }
        for (int k = 0; k < 4; k++) out4[4 * s + k] = acc[k];
    }
}

/* nln/pgrav.cu:713-790 */
int orc_pmm1_mnss_sk1(const float *sink_pos3, int m, const uint16_t *ss_index,
                      const uint32_t *ss_seg, int ss_len, const uint32_t *seg, int seg_len,
                      const float *src_pos3, const int *source_n, double mmass, double eps_inv,
                      double *out4) {
    int bad = 0;
    for (int s = 0; s < m; s++) {
        int si = ss_index[s];
        if (si >= ss_len || ss_seg[2 * si] + ss_seg[2 * si + 1] > (uint32_t)seg_len) {
            bad++;
            continue;
        }
        const uint32_t *sg = seg + 2 * (size_t)ss_seg[2 * si];
        int n = (int)ss_seg[2 * si + 1];
        int cnt = 0;
        for (int g = 0; g < n; g++) cnt += (int)sg[2 * g + 1];
        if (cnt != source_n[si])
            bad++;
        orc_pmm_sk1(sink_pos3 + 3 * s, 1, sg, n, src_pos3, mmass, eps_inv, out4 + 4 * s);
    }
    return bad;
}

/* nln/pgrav.cu:293-368 */
void orc_pq1_mnss(const float *sink_pos3, int m, const uint32_t *seg, const float *ff,
                  int quad_phi_sign, double *out4) {
    for (int s = 0; s < m; s++) {
        const float *f = ff + (size_t)seg[2 * s] * 13;
        orc_pmn(2, 1, sink_pos3 + 3 * s, 1, f, (int)seg[2 * s + 1], 13, 8, 0.0, quad_phi_sign,
                out4 + 4 * s);
    }
}

/* nln/pgrav.cu:498-572 with idx[i] selecting the AoS record (defect A5 fixed) */
void orc_pqq1(const float *sink_pos3, int m, const int *idx, int source_n, const float *qtab13,
              int quad_phi_sign, double *out4) {
    double q[ORC_HSZ];
    for (int s = 0; s < m; s++) {
        double acc[4] = {0, 0, 0, 0};
        for (int i = 0; i < source_n; i++) {
            memset(q, 0, sizeof(q));
            for (int k = 0; k < 13; k++) q[k] = qtab13[(size_t)idx[i] * 13 + k];
            orc_mpole(q, 2, 1, (double)sink_pos3[3 * s] - q[ORC_X],
                      (double)sink_pos3[3 * s + 1] - q[ORC_Y],
                      (double)sink_pos3[3 * s + 2] - q[ORC_Z], 0.0, quad_phi_sign, acc);
        }
        for (int k = 0; k < 4; k++) out4[4 * s + k] = acc[k];
    }
}

/* ---- direct sums (threaded over sinks) ---- */
typedef struct {
    const double *sink, *src, *mass;
    int64_t m0, m1, n;
    int soft_type;
    double eps, L;
    double *out4;
} dsum_job;

static void *dsum_worker(void *arg) {
    dsum_job *j = (dsum_job *)arg;
    for (int64_t s = j->m0; s < j->m1; s++) {
        double acc[4] = {0, 0, 0, 0};
        const double px = j->sink[3 * s], py = j->sink[3 * s + 1], pz = j->sink[3 * s + 2];
        for (int64_t i = 0; i < j->n; i++) {
            double x = px - j->src[3 * i], y = py - j->src[3 * i + 1], z = pz - j->src[3 * i + 2];
            double r2 = x * x + y * y + z * z;
            /* unsoftened: skip the self pair; softened kernels keep it, as the
             * reference's do_gravp (Plummer) and pMM1_mnss_sK1 (K1) do for a
             * body on its own list */
            if (r2 == 0.0 && j->soft_type == 0)
                continue;
            double mi = j->mass[i];
            if (j->soft_type == 4) {
                k1_pair(x, y, z, mi, 1.0 / j->eps, acc);
            } else {
                if (j->soft_type == 1)
                    r2 += j->eps * j->eps;
                double rinv = 1.0 / sqrt(r2);
                double mr = mi * rinv;
                double mr3 = mr * rinv * rinv;
                acc[3] -= mr;
                acc[0] -= x * mr3;
                acc[1] -= y * mr3;
                acc[2] -= z * mr3;
            }
        }
        for (int k = 0; k < 4; k++) j->out4[4 * s + k] = acc[k];
    }
    return NULL;
}

static void *ewald_worker(void *arg) {
    dsum_job *j = (dsum_job *)arg;
    for (int64_t s = j->m0; s < j->m1; s++) {
        double acc[4] = {0, 0, 0, 0};
        for (int64_t i = 0; i < j->n; i++) {
            double x[3] = {j->src[3 * i] - j->sink[3 * s], j->src[3 * i + 1] - j->sink[3 * s + 1],
                           j->src[3 * i + 2] - j->sink[3 * s + 2]};
            double f[3], pot;
            orc_ewald(x, j->L, f, &pot);
            /* nln/grav_n2.c:168-170: q->phi += m*pot; q->acc += m*f */
            acc[0] += j->mass[i] * f[0];
            acc[1] += j->mass[i] * f[1];
            acc[2] += j->mass[i] * f[2];
            acc[3] += j->mass[i] * pot;
        }
        for (int k = 0; k < 4; k++) j->out4[4 * s + k] = acc[k];
    }
    return NULL;
}

static void run_threads(void *(*fn)(void *), dsum_job base, int nthreads) {
    if (nthreads < 1)
        nthreads = 1;
    if (nthreads > 64)
        nthreads = 64;
    pthread_t th[64];
    dsum_job jobs[64];
    int64_t m = base.m1;
    for (int t = 0; t < nthreads; t++) {
        jobs[t] = base;
        jobs[t].m0 = m * t / nthreads;
        jobs[t].m1 = m * (t + 1) / nthreads;
        pthread_create(&th[t], NULL, fn, &jobs[t]);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
}

void orc_direct(const double *sink_pos3, int64_t m, const double *src_pos3,
                const double *src_mass, int64_t n, int soft_type, double eps, double *out4,
                int nthreads) {
    dsum_job b = {sink_pos3, src_pos3, src_mass, 0, m, n, soft_type, eps, 0.0, out4};
    run_threads(dsum_worker, b, nthreads);
}

/* Gradient of L = sum_i w_i . a_i + v_i phi_i for the FP64 direct sum (soft_type
 * 0 none, 1 Plummer), by differentiating every pair in closed form:
 * pair (i <- j), r = x_i - x_j, s^2 = r^2 + eps^2:
 *   a = -m_j r / s^3, phi = -m_j / s
 *   d(w.a + v phi)/dr = m_j ( -w / s^3 + 3 (w.r) r / s^5 + v r / s^3 )
 * added to dL/dx_i and subtracted from dL/dx_j; dL/dm_j += w.(-r/s^3) - v/s. */
typedef struct {
    const double *pos, *mass, *w4;
    int64_t n, k0, k1;
    int soft;
    double eps;
    double *grad4;
} grad_job;

static void *grad_worker(void *arg) {
    grad_job *J = (grad_job *)arg;
    const double e2 = J->soft == 1 ? J->eps * J->eps : 0.0;
    /* body k collects its own sink terms and its source terms (no writes to others) */
    for (int64_t k = J->k0; k < J->k1; k++) {
        double gv[3] = {0, 0, 0}, gm = 0;
        const double *pk = J->pos + 3 * k;
        for (int64_t j = 0; j < J->n; j++) {
            const double *pj = J->pos + 3 * j;
            double r[3] = {pk[0] - pj[0], pk[1] - pj[1], pk[2] - pj[2]};
            double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            if (r2 == 0.0 && J->soft == 0)
                continue;
            double s2 = r2 + e2, s = sqrt(s2), s3 = s2 * s, s5 = s3 * s2;
            /* k as sink, j as source: d/dx_k */
            const double *w = J->w4 + 4 * k;
            double wr = w[0] * r[0] + w[1] * r[1] + w[2] * r[2];
            for (int q = 0; q < 3; q++)
                gv[q] += J->mass[j] * (-w[q] / s3 + 3.0 * wr * r[q] / s5 + w[3] * r[q] / s3);
            /* j as sink, k as source: r' = x_j - x_k = -r; d/dx_k = -d/dr' */
            const double *wj = J->w4 + 4 * j;
            double wr2 = -(wj[0] * r[0] + wj[1] * r[1] + wj[2] * r[2]);
            for (int q = 0; q < 3; q++)
                gv[q] -= J->mass[k] * (-wj[q] / s3 + 3.0 * wr2 * (-r[q]) / s5 + wj[3] * (-r[q]) / s3);
            /* dL/dm_k from body j's result */
            gm += (wj[0] * r[0] + wj[1] * r[1] + wj[2] * r[2]) / s3 - wj[3] / s;
        }
        J->grad4[4 * k + 0] = gv[0];
        J->grad4[4 * k + 1] = gv[1];
        J->grad4[4 * k + 2] = gv[2];
        J->grad4[4 * k + 3] = gm;
    }
    return NULL;
}

void orc_direct_grad(const double *pos3, const double *mass, int64_t n, const double *w4,
                     int soft_type, double eps, double *grad4, int nthreads) {
    if (nthreads < 1)
        nthreads = 1;
    if (nthreads > 64)
        nthreads = 64;
    pthread_t th[64];
    grad_job jobs[64];
    for (int t = 0; t < nthreads; t++) {
        grad_job j = {pos3, mass, w4, n, n * t / nthreads, n * (t + 1) / nthreads, soft_type, eps,
                      grad4};
        jobs[t] = j;
        pthread_create(&th[t], NULL, grad_worker, &jobs[t]);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
}

/* nln/ewald.c:13-64 (alpha = 2/L, |n| <= 4 within 4L, |h|^2 <= 16) */
void orc_ewald(const double x[3], double L, double f[3], double *phi) {
    const int N = 4;
    double alpha = 2.0 / L;
    f[0] = f[1] = f[2] = 0.0;
    *phi = M_PI / (alpha * alpha * L * L * L);
    if (x[0] == 0.0 && x[1] == 0.0 && x[2] == 0.0)
        return;
    double a1 = 2.0 * alpha / sqrt(M_PI);
    double a2 = -M_PI * M_PI / (alpha * alpha * L * L);
    double a3 = 2.0 * M_PI / L;
    double a4 = 2.0 / (L * L);
    double a5 = 1.0 / (M_PI * L);
    for (int i = -N; i <= N; i++)
        for (int j = -N; j <= N; j++)
            for (int k = -N; k <= N; k++) {
                double xx[3] = {x[0] - L * i, x[1] - L * j, x[2] - L * k};
                double ax2 = xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2];
                double axx = sqrt(ax2);
                if (axx > 4.0 * L)
                    continue;
                double t1 = erfc(alpha * axx);
                double t2 = a1 * exp(-alpha * alpha * ax2);
                double t = (t1 / axx + t2) / ax2;
                f[0] += t * xx[0];
                f[1] += t * xx[1];
                f[2] += t * xx[2];
                *phi -= t1 / axx;
            }
    for (int i = -N; i <= N; i++)
        for (int j = -N; j <= N; j++)
            for (int k = -N; k <= N; k++) {
                double hh = (double)(i * i + j * j + k * k);
                if (hh > 16 || hh == 0)
                    continue;
                double hx = i * x[0] + j * x[1] + k * x[2];
                double t = a4 * exp(a2 * hh) * sin(a3 * hx) / hh;
                f[0] += t * i;
                f[1] += t * j;
                f[2] += t * k;
                t = a5 * exp(a2 * hh) * cos(a3 * hx) / hh;
                *phi -= t;
            }
}

void orc_direct_ewald(const double *sink_pos3, int64_t m, const double *src_pos3,
                      const double *src_mass, int64_t n, double L, double *out4, int nthreads) {
    dsum_job b = {sink_pos3, src_pos3, src_mass, 0, m, n, 0, 0.0, L, out4};
    run_threads(ewald_worker, b, nthreads);
}

/* Value and derivatives of the Ewald correction at x = source - sink: the sums
 * of nln/ewald.c:13-64 and their derivatives term by term.
 * Real space, y = x - nL, rho = |y|: t1 = erfc(alpha rho), t2 = a1 exp(-alpha^2 rho^2),
 *   g = t1/rho + t2, t = g/rho^2: f += t y, phi -= t1/rho;
 *   dt1/drho = -t2, dt2/drho = -2 alpha^2 rho t2, so
 *   g' = -t2/rho - t1/rho^2 - 2 alpha^2 rho t2,
 *   g'' = 2 t2/rho^2 + 2 t1/rho^3 + 4 alpha^4 rho^2 t2,
 *   t' = g'/rho^2 - 2 g/rho^3, t'' = g''/rho^2 - 4 g'/rho^3 + 6 g/rho^4;
 *   with u = t'/rho, u' = t''/rho - t'/rho^2:
 *   df_i/dx_j += t delta_ij + u y_i y_j,
 *   d2f_i/dx_j dx_k += u (delta_ij y_k + delta_ik y_j + delta_jk y_i) + (u'/rho) y_i y_j y_k.
 * Fourier space, e = exp(a2 |h|^2), s = sin(a3 h.x), c = cos(a3 h.x):
 *   f += a4 e s/|h|^2 h, phi -= a5 e c/|h|^2;
 *   df_i/dx_j += a3 a4 e c/|h|^2 h_i h_j, d2f_i/dx_j dx_k -= a3^2 a4 e s/|h|^2 h_i h_j h_k.
 * Newtonian term removed: C = f - x/|x|^3, dC_i/dx_j = df_i/dx_j - delta_ij/|x|^3
 *   + 3 x_i x_j/|x|^5, d2C_i/dx_j dx_k = d2f_i/dx_j dx_k + 3 (delta_ij x_k + delta_ik
 *   x_j + delta_jk x_i)/|x|^5 - 15 x_i x_j x_k/|x|^7; psi = phi + 1/|x|, dpsi/dx = C.
 * At x = 0 the n = 0 term minus the Newtonian one is replaced by its limit:
 *   erf(alpha rho)/rho -> a1 in psi; (erfc(alpha rho) - 1)/rho^3 + t2/rho^2 ->
 *   -(4/3) alpha^3/sqrt(pi) = -(2/3) alpha^2 a1 on the diagonal of dC/dx; 0 in the
 *   (odd) second derivatives.
 * T (the second derivatives of C, fully symmetric as third derivatives of psi)
 * is stored as its 10 components xxx, xxy, xxz, xyy, xyz, xzz, yyy, yyz, yzz, zzz. */
static const int SYM3[10][3] = {{0, 0, 0}, {0, 0, 1}, {0, 0, 2}, {0, 1, 1}, {0, 1, 2},
                                {0, 2, 2}, {1, 1, 1}, {1, 1, 2}, {1, 2, 2}, {2, 2, 2}};

static void ewald_value_derivs(const double x[3], double L, double v[4], double J[12],
                               double T[10]) {
    const int N = 4;
    const double alpha = 2.0 / L;
    const double a1 = 2.0 * alpha / sqrt(M_PI);
    const double a2 = -M_PI * M_PI / (alpha * alpha * L * L);
    const double a3 = 2.0 * M_PI / L;
    const double a4 = 2.0 / (L * L);
    const double a5 = 1.0 / (M_PI * L);
    const int origin = x[0] == 0.0 && x[1] == 0.0 && x[2] == 0.0;
    double f[3] = {0, 0, 0}, D[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double H[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    double phi = M_PI / (alpha * alpha * L * L * L);
    for (int i = -N; i <= N; i++)
        for (int j = -N; j <= N; j++)
            for (int k = -N; k <= N; k++) {
                if (origin && i == 0 && j == 0 && k == 0) {
                    phi += a1;
                    for (int a = 0; a < 3; a++) D[a][a] -= 2.0 / 3.0 * alpha * alpha * a1;
                    continue;
                }
                const double y[3] = {x[0] - L * i, x[1] - L * j, x[2] - L * k};
                const double rho2 = y[0] * y[0] + y[1] * y[1] + y[2] * y[2];
                const double rho = sqrt(rho2);
                if (rho > 4.0 * L)
                    continue;
                const double t1 = erfc(alpha * rho);
                const double t2 = a1 * exp(-alpha * alpha * rho2);
                const double g = t1 / rho + t2;
                const double t = g / rho2;
                const double gp = -t2 / rho - t1 / rho2 - 2.0 * alpha * alpha * rho * t2;
                const double gpp = 2.0 * t2 / rho2 + 2.0 * t1 / (rho2 * rho) +
                                   4.0 * alpha * alpha * alpha * alpha * rho2 * t2;
                const double tp = gp / rho2 - 2.0 * g / (rho2 * rho);
                const double tpp = gpp / rho2 - 4.0 * gp / (rho2 * rho) + 6.0 * g / (rho2 * rho2);
                const double u = tp / rho, up = tpp / rho - tp / rho2;
                for (int a = 0; a < 3; a++) {
                    f[a] += t * y[a];
                    D[a][a] += t;
                    for (int b = 0; b < 3; b++) D[a][b] += u * y[a] * y[b];
                }
                for (int q = 0; q < 10; q++) {
                    const int a = SYM3[q][0], b = SYM3[q][1], c = SYM3[q][2];
                    H[q] += u * ((a == b) * y[c] + (a == c) * y[b] + (b == c) * y[a]) +
                            up / rho * y[a] * y[b] * y[c];
                }
                phi -= t1 / rho;
            }
    for (int i = -N; i <= N; i++)
        for (int j = -N; j <= N; j++)
            for (int k = -N; k <= N; k++) {
                const double hh = (double)(i * i + j * j + k * k);
                if (hh > 16 || hh == 0)
                    continue;
                const double h[3] = {(double)i, (double)j, (double)k};
                const double hx = i * x[0] + j * x[1] + k * x[2];
                const double e = exp(a2 * hh);
                const double s = sin(a3 * hx), c = cos(a3 * hx);
                for (int a = 0; a < 3; a++) {
                    f[a] += a4 * e * s / hh * h[a];
                    for (int b = 0; b < 3; b++) D[a][b] += a3 * a4 * e * c / hh * h[a] * h[b];
                }
                for (int q = 0; q < 10; q++)
                    H[q] -= a3 * a3 * a4 * e * s / hh * h[SYM3[q][0]] * h[SYM3[q][1]] * h[SYM3[q][2]];
                phi -= a5 * e * c / hh;
            }
    if (!origin) {
        const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
        const double r = sqrt(r2), r3 = r * r2, r5 = r3 * r2, r7 = r5 * r2;
        for (int a = 0; a < 3; a++) {
            f[a] -= x[a] / r3;
            D[a][a] -= 1.0 / r3;
            for (int b = 0; b < 3; b++) D[a][b] += 3.0 * x[a] * x[b] / r5;
        }
        for (int q = 0; q < 10; q++) {
            const int a = SYM3[q][0], b = SYM3[q][1], c = SYM3[q][2];
            H[q] += 3.0 * ((a == b) * x[c] + (a == c) * x[b] + (b == c) * x[a]) / r5 -
                    15.0 * x[a] * x[b] * x[c] / r7;
        }
        phi += 1.0 / r;
    }
    for (int a = 0; a < 3; a++) {
        v[a] = f[a];
        for (int b = 0; b < 3; b++) J[3 * a + b] = D[a][b];
        J[9 + a] = f[a];
    }
    v[3] = phi;
    for (int q = 0; q < 10; q++) T[q] = H[q];
}

/* index of the symmetric component (a, b, c) in the 10-component order */
static int sym3_index(int a, int b, int c) {
    int s[3] = {a, b, c};
    for (int p = 0; p < 2; p++)
        for (int q = 0; q < 2 - p; q++)
            if (s[q] > s[q + 1]) {
                const int t = s[q];
                s[q] = s[q + 1];
                s[q + 1] = t;
            }
    for (int q = 0; q < 10; q++)
        if (SYM3[q][0] == s[0] && SYM3[q][1] == s[1] && SYM3[q][2] == s[2])
            return q;
    return 0;
}

typedef struct {
    const double *r3;
    int64_t i0, i1;
    double L;
    int nodes;
    double *val4, *jac12, *third10;
} ecorr_job;

static void *ecorr_worker(void *arg) {
    const ecorr_job *j = (const ecorr_job *)arg;
    for (int64_t p = j->i0; p < j->i1; p++) {
        const double *r = j->r3 + 3 * p;
        double *v = j->val4 + 4 * p, *J = j->jac12 + 12 * p;
        double Tn[10];
        double *T = j->third10 ? j->third10 + 10 * p : Tn;
        if (j->nodes <= 0) {
            ewald_value_derivs(r, j->L, v, J, T);
            continue;
        }
        const double h = 0.5 * j->L / (j->nodes - 1);
        double node[3], d[3];
        for (int a = 0; a < 3; a++) {
            double q = rint(r[a] / h);
            if (q > j->nodes - 1)
                q = j->nodes - 1;
            if (q < -(j->nodes - 1))
                q = -(j->nodes - 1);
            node[a] = q * h;
            d[a] = r[a] - node[a];
        }
        double vn[4], Jn[12];
        ewald_value_derivs(node, j->L, vn, Jn, T);
        /* Taylor steps from the node: T d, T d d, T d d d */
        double Td[3][3], Tdd[3], Tddd = 0;
        for (int a = 0; a < 3; a++) {
            Tdd[a] = 0;
            for (int b = 0; b < 3; b++) {
                Td[a][b] = 0;
                for (int c = 0; c < 3; c++) Td[a][b] += T[sym3_index(a, b, c)] * d[c];
                Tdd[a] += Td[a][b] * d[b];
            }
            Tddd += Tdd[a] * d[a];
        }
        double dJd = 0;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) dJd += d[a] * Jn[3 * a + b] * d[b];
        /* psi + C.d + d.J.d / 2 + T d d d / 6 */
        v[3] = vn[3] + vn[0] * d[0] + vn[1] * d[1] + vn[2] * d[2] + 0.5 * dJd + Tddd / 6.0;
        for (int a = 0; a < 3; a++) {
            /* C + J d + T d d / 2 */
            v[a] = vn[a] + Jn[3 * a] * d[0] + Jn[3 * a + 1] * d[1] + Jn[3 * a + 2] * d[2] +
                   0.5 * Tdd[a];
            /* J + T d */
            for (int b = 0; b < 3; b++) J[3 * a + b] = Jn[3 * a + b] + Td[a][b];
        }
        for (int a = 0; a < 3; a++) J[9 + a] = v[a];
    }
    return NULL;
}

void orc_ewald_corr(const double *r3, int64_t n, double L, int nodes, double *val4,
                    double *jac12, double *third10, int nthreads) {
    if (nthreads < 1)
        nthreads = 1;
    if (nthreads > 64)
        nthreads = 64;
    pthread_t th[64];
    ecorr_job jobs[64];
    for (int t = 0; t < nthreads; t++) {
        ecorr_job j = {r3, n * t / nthreads, n * (t + 1) / nthreads, L, nodes, val4, jac12, third10};
        jobs[t] = j;
        pthread_create(&th[t], NULL, ecorr_worker, &jobs[t]);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
}

/* nln/cofm.c:650-720 (mpole_add_mono), summed over bodies about `center` */
void orc_raw_moments(const double *pos3, const double *mass, int64_t n, const double center[3],
                     double raw[35]) {
    memset(raw, 0, 35 * sizeof(double));
    for (int64_t i = 0; i < n; i++) {
        double m = mass[i];
        double x = pos3[3 * i] - center[0], y = pos3[3 * i + 1] - center[1],
               z = pos3[3 * i + 2] - center[2];
        double v[35] = {1,
                        x, y, z,
                        x * x, x * y, x * z, y * y, y * z, z * z,
                        x * x * x, x * x * y, x * x * z, x * y * y, x * y * z, x * z * z,
                        y * y * y, y * y * z, y * z * z, z * z * z,
                        x * x * x * x, x * x * x * y, x * x * x * z, x * x * y * y, x * x * y * z,
                        x * x * z * z, x * y * y * y, x * y * y * z, x * y * z * z, x * z * z * z,
                        y * y * y * y, y * y * y * z, y * y * z * z, y * z * z * z, z * z * z * z};
        for (int k = 0; k < 35; k++) raw[k] += m * v[k];
    }
}

/* nln/cofm.c:398-562 (CellFromCofm, hexacell branch), FP64 */
void orc_cell_record(const double *pos3, const double *mass, int64_t n, const double center[3],
                     double bmax, double rec[ORC_HSZ]) {
    double u[35];
    orc_raw_moments(pos3, mass, n, center, u);
    enum { M, X, Y, Z, X2, XY, XZ, Y2, YZ, Z2, X3, X2Y, X2Z, XY2, XYZ, XZ2, Y3, Y2Z, YZ2, Z3,
           X4, X3Y, X3Z, X2Y2, X2YZ, X2Z2, XY3, XY2Z, XYZ2, XZ3, Y4, Y3Z, Y2Z2, YZ3, Z4 };
    double rinv = 1.0 / bmax;
    rec[ORC_MASS] = u[M];
    rec[ORC_X] = center[0];
    rec[ORC_Y] = center[1];
    rec[ORC_Z] = center[2];
    rec[ORC_R] = bmax;
    rec[ORC_QX] = rinv * u[X];
    rec[ORC_QY] = rinv * u[Y];
    rec[ORC_QZ] = rinv * u[Z];
    double t = (u[X2] + u[Y2] + u[Z2]) / 3.0;
    double r2 = rinv * rinv;
    rec[ORC_QXX] = (u[X2] - t) * r2;
    rec[ORC_QXY] = u[XY] * r2;
    rec[ORC_QYY] = (u[Y2] - t) * r2;
    rec[ORC_QXZ] = u[XZ] * r2;
    rec[ORC_QYZ] = u[YZ] * r2;
    double tx = (u[X3] + u[XY2] + u[XZ2]) / 5.0;
    double ty = (u[X2Y] + u[Y3] + u[YZ2]) / 5.0;
    double tz = (u[X2Z] + u[Y2Z] + u[Z3]) / 5.0;
    double r3 = r2 * rinv;
    rec[ORC_QXXX] = (u[X3] - 3.0 * tx) * r3;
    rec[ORC_QXXY] = (u[X2Y] - ty) * r3;
    rec[ORC_QXYY] = (u[XY2] - tx) * r3;
    rec[ORC_QYYY] = (u[Y3] - 3.0 * ty) * r3;
    rec[ORC_QXXZ] = (u[X2Z] - tz) * r3;
    rec[ORC_QXYZ] = u[XYZ] * r3;
    rec[ORC_QYYZ] = (u[Y2Z] - tz) * r3;
    double txx = (u[X4] + u[X2Y2] + u[X2Z2]) / 7.0;
    double txy = (u[X3Y] + u[XY3] + u[XYZ2]) / 7.0;
    double txz = (u[X3Z] + u[XY2Z] + u[XZ3]) / 7.0;
    double tyy = (u[X2Y2] + u[Y4] + u[Y2Z2]) / 7.0;
    double tyz = (u[X2YZ] + u[Y3Z] + u[YZ3]) / 7.0;
    double tzz = (u[X2Z2] + u[Y2Z2] + u[Z4]) / 7.0;
    t = 0.1 * (txx + tyy + tzz);
    double r4 = r2 * r2;
    rec[ORC_QXXXX] = (u[X4] - 6.0 * (txx - t)) * r4;
    rec[ORC_QXXXY] = (u[X3Y] - 3.0 * txy) * r4;
    rec[ORC_QXXYY] = (u[X2Y2] - (txx + tyy - 2.0 * t)) * r4;
    rec[ORC_QXYYY] = (u[XY3] - 3.0 * txy) * r4;
    rec[ORC_QYYYY] = (u[Y4] - 6.0 * (tyy - t)) * r4;
    rec[ORC_QXXXZ] = (u[X3Z] - 3.0 * txz) * r4;
    rec[ORC_QXXYZ] = (u[X2YZ] - tyz) * r4;
    rec[ORC_QXYYZ] = (u[XY2Z] - txz) * r4;
    rec[ORC_QYYYZ] = (u[Y3Z] - 3.0 * tyz) * r4;
}

/* ------------------------------------------------------------ integrator */
#define FLD(p, i, stride) ((const float *)((const char *)(p) + (size_t)(i) * (size_t)(stride)))
#define OFLD(p, i, stride) ((float *)((char *)(p) + (size_t)(i) * (size_t)(stride)))

/* nln/integrate.c:16-59 */
void orc_integrate(const float *inmass, const float *inpos, const float *invel, const float *inacc,
                   const float *inphi, int instride, float *outpos, float *outvel, int outstride,
                   int n, double dt, double dtv, double *tpos, double *tvel, double *kep,
                   double *pep) {
    const double dtsync = *tpos - *tvel; /* :32 */
    double ke = 0.0, pe = 0.0;
    for (int i = 0; i < n; i++) {
        const float m = *FLD(inmass, i, instride), phi = *FLD(inphi, i, instride);
        const float *x = FLD(inpos, i, instride), *v = FLD(invel, i, instride),
                    *a = FLD(inacc, i, instride);
        float *ox = OFLD(outpos, i, outstride), *ov = OFLD(outvel, i, outstride);
        double vc[3];
        float nv[3], nx[3];
        for (int k = 0; k < 3; k++) vc[k] = v[k] + dtsync * a[k];    /* :40 */
        ke += m * (vc[0] * vc[0] + vc[1] * vc[1] + vc[2] * vc[2]);  /* :41 */
        pe += m * phi;                                              /* :42, a float product */
        for (int k = 0; k < 3; k++) nv[k] = (float)(v[k] + dtv * a[k]);   /* :43 */
        for (int k = 0; k < 3; k++) nx[k] = (float)(x[k] + dt * nv[k]);   /* :44 */
        for (int k = 0; k < 3; k++) {
            ov[k] = nv[k];
            ox[k] = nx[k];
        }
    }
    *tvel += dtv; /* :54-57 */
    *tpos += dt;
    *kep = 0.5 * ke;
    *pep = 0.5 * pe;
}

/* nln/integrate.c:82-135, the cosmology's factors given */
void orc_cosmo_integrate(const float *inmass, const float *inpos, const float *invel,
                         const float *inacc, const float *inphi, int instride, float *outpos,
                         float *outvel, int outstride, int n, double a0, double a1,
                         double dt_kick_topos, double dt_kick, double dt_drift, double *kep,
                         double *pep) {
    double ke = 0.0, pe = 0.0;
    for (int i = 0; i < n; i++) {
        const float m = *FLD(inmass, i, instride), phi = *FLD(inphi, i, instride);
        const float *x = FLD(inpos, i, instride), *v = FLD(invel, i, instride),
                    *a = FLD(inacc, i, instride);
        float *ox = OFLD(outpos, i, outstride), *ov = OFLD(outvel, i, outstride);
        double vc[3];
        float nv[3], nx[3];
        for (int k = 0; k < 3; k++) vc[k] = v[k] + dt_kick_topos * a[k];              /* :115 */
        ke += m * (vc[0] * vc[0] + vc[1] * vc[1] + vc[2] * vc[2]) / (a0 * a0);       /* :116 */
        pe += m * phi;                                                              /* :117 */
        for (int k = 0; k < 3; k++) nv[k] = (float)(v[k] + dt_kick * a[k]);           /* :118 */
        for (int k = 0; k < 3; k++) nx[k] = (float)((1.0 / a0) * x[k]);               /* :119 */
        for (int k = 0; k < 3; k++) nx[k] = (float)(nx[k] + dt_drift * nv[k]);        /* :120 */
        for (int k = 0; k < 3; k++) nx[k] = (float)(a1 * nx[k]);                      /* :121 */
        for (int k = 0; k < 3; k++) {
            ov[k] = nv[k];
            ox[k] = nx[k];
        }
    }
    *kep = 0.5 * ke;
    *pep = 0.5 * pe;
}

/* nln/integrate.c:62-80 */
void orc_convert_v(float *vel, int stride, int n, double scale, int to_physical) {
    const float one_on_scale = (float)1.0 / scale;
    for (int i = 0; i < n; i++) {
        float *v = OFLD(vel, i, stride);
        for (int k = 0; k < 3; k++) {
            if (to_physical)
                v[k] *= one_on_scale;
            else
                v[k] = (float)(v[k] * scale);
        }
    }
}
