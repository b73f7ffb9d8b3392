/*
 * TEST INFRASTRUCTURE ONLY -- CPU oracle of the SMA law (stress model 160).
 *
 * Restates, in plain C and FP64, the arithmetic of
 *   sld_stress_model_160.f90 (eval_sld_stress_model_160 :36-195 with
 *     flagt=0, compute_kirchoff_stress :212-500, eval_loading_function
 *     :513-537, eval_beta :539-556, INV2x2 :558-573, INV3x3 :575-599,
 *     OUTPROD_v3 :601-614),
 *   dsyevq3.f90:22-137, dsytrd3.f90:22-109 (Kopp's Householder + QL),
 *   eigsrt.f90:19-48 (descending selection sort despite its comment).
 * The order of every floating-point operation follows the Fortran source so
 * that, compiled with -ffp-contract=off, results agree with the compiled
 * reference to rounding.  Quirks kept on purpose:
 *   - temperature is inert: gp_tmp = 0 (:105);
 *   - beta is forced to 1 (:554);
 *   - the first Newton residual is squared, later ones are L2 (:403, :454);
 *   - xS accumulates the running dh totals every iteration (:429-432), so
 *     xS != xS_0 + dh_AS + dh_SA after the first iteration;
 *   - on 65 iterations the history reverts to the TIME_N values (:459-465);
 *   - F_AS/F_SA are stored before the [0,1] clamp of xS (:484, :496-498).
 */
#include <math.h>
#include <float.h>
#include <string.h>
#include "oracle.h"

/* column-major 3x3 access: A(i,j), 0-based */
#define M3(A, i, j) (A)[(i) + 3 * (j)]

/* dsytrd3.f90:22-109 -- Householder reduction to tridiagonal form. */
static void dsytrd3(const double *A, double *Q, double d[3], double e[3])
{
    double u[3], p[3], omega, f, K, h, g;
    int i, j;
    for (i = 0; i < 3; i++) {
        M3(Q, i, i) = 1.0;
        for (j = 0; j < i; j++) { M3(Q, i, j) = 0.0; M3(Q, j, i) = 0.0; }
    }
    h = M3(A, 0, 1) * M3(A, 0, 1) + M3(A, 0, 2) * M3(A, 0, 2);
    if (M3(A, 0, 1) > 0.0) g = -sqrt(h); else g = sqrt(h);
    e[0] = g;
    f = g * M3(A, 0, 1);
    u[1] = M3(A, 0, 1) - g;
    u[2] = M3(A, 0, 2);
    omega = h - f;
    if (omega > 0.0) {
        omega = 1.0 / omega;
        K = 0.0;
        for (i = 1; i < 3; i++) {
            f = M3(A, 1, i) * u[1] + M3(A, i, 2) * u[2];
            p[i] = omega * f;
            K = K + u[i] * f;
        }
        K = 0.5 * K * (omega * omega);
        for (i = 1; i < 3; i++) p[i] = p[i] - K * u[i];
        d[0] = M3(A, 0, 0);
        d[1] = M3(A, 1, 1) - 2.0 * p[1] * u[1];
        d[2] = M3(A, 2, 2) - 2.0 * p[2] * u[2];
        for (j = 1; j < 3; j++) {
            f = omega * u[j];
            for (i = 1; i < 3; i++) M3(Q, i, j) = M3(Q, i, j) - f * u[i];
        }
        e[1] = M3(A, 1, 2) - p[1] * u[2] - u[1] * p[2];
    } else {
        for (i = 0; i < 3; i++) d[i] = M3(A, i, i);
        e[1] = M3(A, 1, 2);
    }
}

/* dsyevq3.f90:22-137 -- QL with implicit shifts.  Returns 1 when the
 * 30-iteration bail-out (:82-84) fires, leaving Q and w as they are. */
int orc_dsyevq3(const double *A, double *Q, double w[3])
{
    double e[3] = {0.0, 0.0, 0.0};
    double g, r, p, f, b, s, c, t;
    int l, m, i, j, k, niter;
    dsytrd3(A, Q, w, e);
    for (l = 0; l < 2; l++) {
        niter = 0;
        for (i = 0; i < 50; i++) {
            for (m = l; m < 2; m++) {
                g = fabs(w[m]) + fabs(w[m + 1]);
                if (fabs(e[m]) + g == g) break;
            }
            if (m == l) break;
            niter = niter + 1;
            if (niter >= 30) return 1;
            g = (w[l + 1] - w[l]) / (2.0 * e[l]);
            r = sqrt(1.0 + g * g);
            if (g >= 0.0) g = w[m] - w[l] + e[l] / (g + r);
            else          g = w[m] - w[l] + e[l] / (g - r);
            s = 1.0; c = 1.0; p = 0.0;
            for (j = m - 1; j >= l; j--) {
                f = s * e[j];
                b = c * e[j];
                if (fabs(f) > fabs(g)) {
                    c = g / f;
                    r = sqrt(1.0 + c * c);
                    e[j + 1] = f * r;
                    s = 1.0 / r;
                    c = c * s;
                } else {
                    s = f / g;
                    r = sqrt(1.0 + s * s);
                    e[j + 1] = g * r;
                    c = 1.0 / r;
                    s = s * c;
                }
                g = w[j + 1] - p;
                r = (w[j] - g) * s + 2.0 * c * b;
                p = s * r;
                w[j + 1] = g + p;
                g = c * r - b;
                for (k = 0; k < 3; k++) {
                    t = M3(Q, k, j + 1);
                    M3(Q, k, j + 1) = s * M3(Q, k, j) + c * t;
                    M3(Q, k, j) = c * M3(Q, k, j) - s * t;
                }
            }
            w[l] = w[l] - p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    return 0;
}

/* eigsrt.f90:19-48 -- selection sort, largest first ('>=' keeps the last
 * of equal values), eigenvector columns swapped alongside. */
void orc_eigsrt(double d[3], double *V)
{
    int i, j, k;
    double p;
    for (i = 0; i < 2; i++) {
        k = i;
        p = d[i];
        for (j = i + 1; j < 3; j++) {
            if (d[j] >= p) { k = j; p = d[j]; }
        }
        if (k != i) {
            d[k] = d[i];
            d[i] = p;
            for (j = 0; j < 3; j++) {
                p = M3(V, j, i);
                M3(V, j, i) = M3(V, j, k);
                M3(V, j, k) = p;
            }
        }
    }
}

/* eval_loading_function :513-537 with eval_beta :539-556 (beta = 1). */
static double loading_function(double K, double G, double alfa, double C, double eL,
                               double theta, double enor, double tmp, double xS)
{
    double pres = K * (theta - 3.0 * alfa * eL * xS);
    double trac = 2.0 * G * (enor - eL * xS);
    double beta = 1.0;
    return trac * beta + 3.0 * alfa * pres - C * tmp;
}

/* INV3x3 :575-599 (division per entry, as in the reference). */
static double inv3x3(const double *M, double *Mi)
{
    double det = M3(M,0,0)*M3(M,1,1)*M3(M,2,2) + M3(M,0,2)*M3(M,1,0)*M3(M,2,1)
               + M3(M,2,0)*M3(M,0,1)*M3(M,1,2) - M3(M,2,0)*M3(M,1,1)*M3(M,0,2)
               - M3(M,2,2)*M3(M,0,1)*M3(M,1,0) - M3(M,0,0)*M3(M,1,2)*M3(M,2,1);
    M3(Mi,0,0) =  (M3(M,1,1)*M3(M,2,2) - M3(M,2,1)*M3(M,1,2))/det;
    M3(Mi,0,1) = -(M3(M,0,1)*M3(M,2,2) - M3(M,0,2)*M3(M,2,1))/det;
    M3(Mi,0,2) =  (M3(M,0,1)*M3(M,1,2) - M3(M,1,1)*M3(M,0,2))/det;
    M3(Mi,1,0) = -(M3(M,1,0)*M3(M,2,2) - M3(M,2,0)*M3(M,1,2))/det;
    M3(Mi,1,1) =  (M3(M,0,0)*M3(M,2,2) - M3(M,0,2)*M3(M,2,0))/det;
    M3(Mi,1,2) = -(M3(M,0,0)*M3(M,1,2) - M3(M,1,0)*M3(M,0,2))/det;
    M3(Mi,2,0) =  (M3(M,1,0)*M3(M,2,1) - M3(M,2,0)*M3(M,1,1))/det;
    M3(Mi,2,1) = -(M3(M,0,0)*M3(M,2,1) - M3(M,2,0)*M3(M,0,1))/det;
    M3(Mi,2,2) =  (M3(M,0,0)*M3(M,1,1) - M3(M,0,1)*M3(M,1,0))/det;
    return det;
}

/* compute_kirchoff_stress :212-500.  Returns status bits. */
static uint32_t kirchhoff(const double *props, const double *Fg, double tmp,
                          double SDV[3], double *STR)
{
    const int iter_NR = 65;
    const double tole_NR = 1.0e-8;
    uint32_t flags = 0;
    double E = props[0], v = props[1], eL = props[2];
    double Ss_AS = props[3], Sf_AS = props[4], C_AS = props[5];
    double Tf_AS = props[6], Ts_AS = props[7];
    double Ss_SA = props[9], Sf_SA = props[10], C_SA = props[11];
    double Tf_SA = props[12], Ts_SA = props[13], alfa = props[14];
    double K = E / (3.0 - 6.0 * v);
    double G = E / (2.0 + 2.0 * v);
    double aux = sqrt(2.0 / 3.0) + alfa;
    double Rf_AS = Sf_AS * aux + C_AS * Tf_AS;
    double Rs_AS = Ss_AS * aux + C_AS * Ts_AS;
    double Rf_SA = Sf_SA * aux + C_SA * Tf_SA;
    double Rs_SA = Ss_SA * aux + C_SA * Ts_SA;
    double G_1 = (2.0 * G + 9.0 * (alfa * alfa) * K) * eL;
    double b[9], nA[9], hA[3];
    int i, j, k;

    /* b = F F^T (matmul order: sum over k ascending) */
    for (j = 0; j < 3; j++)
        for (i = 0; i < 3; i++) {
            double s = 0.0;
            for (k = 0; k < 3; k++) s = s + M3(Fg, i, k) * M3(Fg, j, k);
            M3(b, i, j) = s;
        }
    if (orc_dsyevq3(b, nA, hA)) flags |= ORC_FLAG_EIGEN_FAIL;
    orc_eigsrt(hA, nA);
    double h_1 = sqrt(hA[0]), h_2 = sqrt(hA[1]), h_3 = sqrt(hA[2]);

    double J_vol = h_1 * h_2 * h_3;
    double theta = log(J_vol);
    double J_dev = pow(J_vol, -1.0 / 3.0);
    double e_1 = log(J_dev * h_1);
    double e_2 = log(J_dev * h_2);
    double e_3 = log(J_dev * h_3);
    double e_n = sqrt(e_1 * e_1 + e_2 * e_2 + e_3 * e_3);
    double d_1, d_2, d_3;
    if (e_n > DBL_EPSILON) { d_1 = e_1 / e_n; d_2 = e_2 / e_n; d_3 = e_3 / e_n; }
    else { d_1 = 0.0; d_2 = 0.0; d_3 = 0.0; }

    double xS_0 = SDV[0], F_AS_0 = SDV[1], F_SA_0 = SDV[2];
    double xS, H_AS, H_SA;
    double F_AS = loading_function(K, G, alfa, C_AS, eL, theta, e_n, tmp, xS_0);
    double Fs_AS = F_AS - Rs_AS;
    double F_AS_dot = F_AS - F_AS_0;
    int kf_AS = (F_AS > F_AS_0 && Fs_AS > 0.0 && xS_0 < 1.0);
    H_AS = kf_AS ? 1.0 : 0.0;
    double F_SA = loading_function(K, G, alfa, C_SA, eL, theta, e_n, tmp, xS_0);
    double Fs_SA = F_SA - Rs_SA;
    double F_SA_dot = F_SA - F_SA_0;
    int kf_SA = (F_SA < F_SA_0 && Fs_SA < 0.0 && xS_0 > 0.0);
    H_SA = kf_SA ? 1.0 : 0.0;
    if (kf_AS) flags |= ORC_FLAG_KF_AS;
    if (kf_SA) flags |= ORC_FLAG_KF_SA;

    if (kf_SA || kf_AS) {
        double dh_SA = 0.0, dh_AS = 0.0, RHS_AS, RHS_SA, resi;
        int iter = 0;
        xS = xS_0;
        RHS_AS = (F_AS - Rf_AS) * dh_AS + H_AS * F_AS_dot * (1.0 - xS);
        RHS_SA = (F_SA - Rf_SA) * dh_SA - H_SA * F_SA_dot * xS;
        resi = RHS_AS * RHS_AS + RHS_SA * RHS_SA;
        while (iter < iter_NR && resi > tole_NR) {
            iter = iter + 1;
            double m11 = -G_1 * dh_AS - (H_AS * F_AS_dot + H_AS * (1.0 - xS) * G_1) + (F_AS - Rf_AS);
            double m12 = -G_1 * dh_AS - (H_AS * F_AS_dot + H_AS * (1.0 - xS) * G_1);
            double m21 = -G_1 * dh_SA - (H_SA * F_SA_dot - H_SA * xS * G_1);
            double m22 = -G_1 * dh_SA - (H_SA * F_SA_dot - H_SA * xS * G_1) + (F_SA - Rf_SA);
            /* INV2x2 :558-573 then matmul(R, rhs) */
            double det = m11 * m22 - m12 * m21;
            double r11 = m22 / det, r12 = -m12 / det, r21 = -m21 / det, r22 = m11 / det;
            double sol1 = r11 * RHS_AS + r12 * RHS_SA;
            double sol2 = r21 * RHS_AS + r22 * RHS_SA;
            dh_AS = dh_AS - sol1;
            dh_SA = dh_SA - sol2;
            xS = xS + dh_SA + dh_AS;
            F_AS = loading_function(K, G, alfa, C_AS, eL, theta, e_n, tmp, xS);
            F_AS_dot = F_AS - F_AS_0;
            F_SA = loading_function(K, G, alfa, C_SA, eL, theta, e_n, tmp, xS);
            F_SA_dot = F_SA - F_SA_0;
            RHS_AS = (F_AS - Rf_AS) * dh_AS + H_AS * F_AS_dot * (1.0 - xS);
            RHS_SA = (F_SA - Rf_SA) * dh_SA - H_SA * F_SA_dot * xS;
            resi = sqrt(RHS_AS * RHS_AS + RHS_SA * RHS_SA);
        }
        if (iter >= iter_NR) {
            flags |= ORC_FLAG_NEWTON_FAIL;
            xS = SDV[0];
            F_AS = SDV[1];
            F_SA = SDV[2];
            xS = xS_0;
        }
        flags |= ((uint32_t)iter & 0xffu) << ORC_FLAG_ITER_SHIFT;
    } else {
        xS = xS_0;
    }
    xS = fmax(fmin(1.0, xS), 0.0);

    /* STR = K(theta-3 alfa eL xS) I + sum_i 2G(e_i - eL xS d_i) n_i n_i^T */
    {
        double cvol = K * (theta - 3.0 * alfa * eL * xS);
        double c1 = 2.0 * G * (e_1 - eL * xS * d_1);
        double c2 = 2.0 * G * (e_2 - eL * xS * d_2);
        double c3 = 2.0 * G * (e_3 - eL * xS * d_3);
        for (j = 0; j < 3; j++)
            for (i = 0; i < 3; i++) {
                double id = (i == j) ? 1.0 : 0.0;
                M3(STR, i, j) = cvol * id
                              + c1 * (M3(nA, j, 0) * M3(nA, i, 0))
                              + c2 * (M3(nA, j, 1) * M3(nA, i, 1))
                              + c3 * (M3(nA, j, 2) * M3(nA, i, 2));
            }
    }
    SDV[0] = xS;
    SDV[1] = F_AS;
    SDV[2] = F_SA;
    return flags;
}

uint32_t orc_sm160_point(const double props[18], const double F[9],
                         const double sdv_old[3], double sdv_new[3],
                         double S[9], double tau[9])
{
    double Finv[9], str[9], sdv[3];
    uint32_t flags = 0;
    int i, j, k, l;
    double det = inv3x3(F, Finv);
    if (!(det > 0.0)) flags |= ORC_FLAG_DETF_NONPOS;
    sdv[0] = sdv_old[0]; sdv[1] = sdv_old[1]; sdv[2] = sdv_old[2];
    flags |= kirchhoff(props, F, 0.0, sdv, str);
    sdv_new[0] = sdv[0]; sdv_new[1] = sdv[1]; sdv_new[2] = sdv[2];
    /* S_ij = sum_k sum_l Finv_ik Finv_jl tau_kl  (:159-168) */
    for (i = 0; i < 3; i++)
        for (j = 0; j < 3; j++) M3(S, i, j) = 0.0;
    for (i = 0; i < 3; i++)
        for (j = 0; j < 3; j++)
            for (k = 0; k < 3; k++)
                for (l = 0; l < 3; l++)
                    M3(S, i, j) = M3(S, i, j) + M3(Finv, i, k) * M3(Finv, j, l) * M3(str, k, l);
    if (tau) memcpy(tau, str, sizeof(str));
    return flags;
}

int orc_sm160_eval(int64_t n, const double *props, const int32_t *mat_of_gp,
                   const double *E_of_gp, const double *F, const double *T,
                   const double *sdv_old, double *sdv_new, double *S,
                   uint32_t *flags)
{
    (void)T; /* temperature is inert in the reference (:105) */
    int64_t i;
    #pragma omp parallel for schedule(dynamic, 256)
    for (i = 0; i < n; i++) {
        double p[18], Fi[9], so[3], sn[3], Si[9];
        int k;
        const double *pm = props + 18 * (mat_of_gp ? mat_of_gp[i] : 0);
        for (k = 0; k < 18; k++) p[k] = pm[k];
        if (E_of_gp) p[0] = E_of_gp[i];
        for (k = 0; k < 9; k++) Fi[k] = F[k * n + i];
        for (k = 0; k < 3; k++) so[k] = sdv_old[k * n + i];
        uint32_t f = orc_sm160_point(p, Fi, so, sn, Si, 0);
        for (k = 0; k < 3; k++) sdv_new[k * n + i] = sn[k];
        for (k = 0; k < 9; k++) S[k * n + i] = Si[k];
        if (flags) flags[i] = f;
    }
    return 0;
}
