/* TEST INFRASTRUCTURE ONLY -- CPU oracle of the neo-Hookean law (Ansys
 * form), restated from sld_stress_model_102.f90 (kfl_serei_sld = 0, the
 * default of sld_reanut.f90:54; modulating factor gpmof = 1):
 *
 *   C = F^T F                                   (:100-103)
 *   C^-1 by invmtx.f90 case(3)                  (:109)
 *   tr C                                        (:111-114)
 *   S = ((-mu0/3) J^(-2/3) tr C + J K0 (J-1)) C^-1 + mu0 J^(-2/3) I   (:116-119)
 *
 * with mu0 = parco(1), K0 = parco(2) (sld_stress_model_102.f90:60-61;
 * NEOHO ANSYS in sld_reaphy.f90:428-431) and J = det F from the element
 * kinematics.  The reference file is not compiled here (its module chain
 * reaches def_domain and mpi_f08); the restatement is pinned by the
 * strain-energy derivative, closed forms and frame indifference in
 * tests/test_oracle_nh102.py.
 */
#include <math.h>
#include "oracle.h"

/* invmtx.f90 case(3), column-major */
static double invmtx3(const double *a, double *b)
{
#define A(i, j) a[(i - 1) + 3 * (j - 1)]
#define B(i, j) b[(i - 1) + 3 * (j - 1)]
    double t1 =  A(2,2) * A(3,3) - A(3,2) * A(2,3);
    double t2 = -A(2,1) * A(3,3) + A(3,1) * A(2,3);
    double t3 =  A(2,1) * A(3,2) - A(3,1) * A(2,2);
    double deter = A(1,1) * t1 + A(1,2) * t2 + A(1,3) * t3;
    if (fabs(deter) == 0.0) return 0.0;
    double denom = 1.0 / deter;
    B(1,1) = t1 * denom;
    B(2,1) = t2 * denom;
    B(3,1) = t3 * denom;
    B(2,2) = ( A(1,1) * A(3,3) - A(3,1) * A(1,3)) * denom;
    B(3,2) = (-A(1,1) * A(3,2) + A(1,2) * A(3,1)) * denom;
    B(3,3) = ( A(1,1) * A(2,2) - A(2,1) * A(1,2)) * denom;
    B(1,2) = (-A(1,2) * A(3,3) + A(3,2) * A(1,3)) * denom;
    B(1,3) = ( A(1,2) * A(2,3) - A(2,2) * A(1,3)) * denom;
    B(2,3) = (-A(1,1) * A(2,3) + A(2,1) * A(1,3)) * denom;
#undef A
#undef B
    return deter;
}

void orc_nh102_point(const double props[18], const double F[9], double J, double S[9])
{
    const double mu0 = props[0], K0 = props[1], gpmof = 1.0;
    double C[9], Ci[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0.0;
            for (int k = 0; k < 3; k++) s = s + F[k + 3 * i] * F[k + 3 * j];
            C[i + 3 * j] = s;
        }
    for (int q = 0; q < 9; q++) Ci[q] = (q % 4 == 0) ? 1.0 : 0.0;   /* gpcin = tkron */
    invmtx3(C, Ci);
    double tracec = 0.0;
    for (int d = 0; d < 3; d++) tracec = tracec + C[4 * d];
    const double jm23 = pow(J, -2.0 / 3.0);
    const double a = ((-mu0 * gpmof) / 3.0) * jm23 * tracec + J * K0 * (J - 1.0);
    const double b = mu0 * gpmof * jm23;
    for (int q = 0; q < 9; q++) S[q] = a * Ci[q] + b * ((q % 4 == 0) ? 1.0 : 0.0);
}
