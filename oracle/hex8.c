/*
 * TEST INFRASTRUCTURE ONLY -- CPU oracle of the HEX08 kinematics.
 *
 *   orc_hex8_tables     mod_lagrange.f90:935-1000 (trilinear brick, node
 *                       order (-,-,-) (+,-,-) (+,+,-) (-,+,-) then z=+1),
 *                       mod_quadrature_gauss.f90:2320-2321 (the 15-digit
 *                       +-0.577350269189626 literal, weights 1) and
 *                       :2372-2385 (xi outer, eta, zeta inner).
 *   orc_hex8_kinematics jacobi.f90 ndime==3 branch (J = elcod deriv^T,
 *                       cofactor inverse, gpcar = J^-T deriv), elmca2.f90:
 *                       105-111 (gpvol = w detJ), sld_elmpre.f90:149-162,
 *                       219-226 (F = I + sum_a u_a (x) gpcar_a; detF by
 *                       invmtx.f90 case 3).
 */
#include <string.h>
#include "oracle.h"

void orc_hex8_tables(double *posgp, double *weigp, double *shapf, double *deriv)
{
    const double posgl[2] = {-0.577350269189626, 0.577350269189626};
    int ig = 0, i, j, k;
    for (i = 0; i < 2; i++)
        for (j = 0; j < 2; j++)
            for (k = 0; k < 2; k++) {
                double s = posgl[i], t = posgl[j], z = posgl[k];
                double sm = 0.5 * (1.0 - s), tm = 0.5 * (1.0 - t), zm = 0.5 * (1.0 - z);
                double sq = 0.5 * (1.0 + s), tp = 0.5 * (1.0 + t), zp = 0.5 * (1.0 + z);
                double *N = shapf + 8 * ig, *D = deriv + 24 * ig;
                posgp[3 * ig + 0] = s; posgp[3 * ig + 1] = t; posgp[3 * ig + 2] = z;
                weigp[ig] = 1.0 * 1.0 * 1.0;
                N[0] = sm * tm * zm; N[1] = sq * tm * zm; N[2] = sq * tp * zm; N[3] = sm * tp * zm;
                N[4] = sm * tm * zp; N[5] = sq * tm * zp; N[6] = sq * tp * zp; N[7] = sm * tp * zp;
                D[0*3+0] = -0.5*tm*zm; D[0*3+1] = -0.5*sm*zm; D[0*3+2] = -0.5*sm*tm;
                D[1*3+0] =  0.5*tm*zm; D[1*3+1] = -0.5*sq*zm; D[1*3+2] = -0.5*sq*tm;
                D[2*3+0] =  0.5*tp*zm; D[2*3+1] =  0.5*sq*zm; D[2*3+2] = -0.5*sq*tp;
                D[3*3+0] = -0.5*tp*zm; D[3*3+1] =  0.5*sm*zm; D[3*3+2] = -0.5*sm*tp;
                D[4*3+0] = -0.5*tm*zp; D[4*3+1] = -0.5*sm*zp; D[4*3+2] =  0.5*sm*tm;
                D[5*3+0] =  0.5*tm*zp; D[5*3+1] = -0.5*sq*zp; D[5*3+2] =  0.5*sq*tm;
                D[6*3+0] =  0.5*tp*zp; D[6*3+1] =  0.5*sq*zp; D[6*3+2] =  0.5*sq*tp;
                D[7*3+0] = -0.5*tp*zp; D[7*3+1] =  0.5*sm*zp; D[7*3+2] =  0.5*sm*tp;
                ig++;
            }
}

/* jacobi.f90 ndime==3: xjacm(i,j) = sum_k elcod(i,k) deriv(j,k). */
static double jacobi3(const double *elcod, const double *D, double *gpcar)
{
    double xm[3][3] = {{0}}, xi[3][3], t1, t2, t3, det, denom;
    int k, j;
    for (k = 0; k < 8; k++) {
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                xm[a][b] = xm[a][b] + elcod[3 * k + a] * D[3 * k + b];
    }
    t1 =  xm[1][1] * xm[2][2] - xm[2][1] * xm[1][2];
    t2 = -xm[1][0] * xm[2][2] + xm[2][0] * xm[1][2];
    t3 =  xm[1][0] * xm[2][1] - xm[2][0] * xm[1][1];
    det = xm[0][0] * t1 + xm[0][1] * t2 + xm[0][2] * t3;
    if (det == 0.0) return det;
    denom = 1.0 / det;
    xi[0][0] = t1 * denom;
    xi[1][0] = t2 * denom;
    xi[2][0] = t3 * denom;
    xi[1][1] = ( xm[0][0] * xm[2][2] - xm[2][0] * xm[0][2]) * denom;
    xi[2][1] = (-xm[0][0] * xm[2][1] + xm[0][1] * xm[2][0]) * denom;
    xi[2][2] = ( xm[0][0] * xm[1][1] - xm[1][0] * xm[0][1]) * denom;
    xi[0][1] = (-xm[0][1] * xm[2][2] + xm[2][1] * xm[0][2]) * denom;
    xi[0][2] = ( xm[0][1] * xm[1][2] - xm[1][1] * xm[0][2]) * denom;
    xi[1][2] = (-xm[0][0] * xm[1][2] + xm[1][0] * xm[0][2]) * denom;
    for (j = 0; j < 8; j++) {
        gpcar[3*j+0] = xi[0][0] * D[3*j+0] + xi[1][0] * D[3*j+1] + xi[2][0] * D[3*j+2];
        gpcar[3*j+1] = xi[0][1] * D[3*j+0] + xi[1][1] * D[3*j+1] + xi[2][1] * D[3*j+2];
        gpcar[3*j+2] = xi[0][2] * D[3*j+0] + xi[1][2] * D[3*j+1] + xi[2][2] * D[3*j+2];
    }
    return det;
}

/* invmtx.f90 case 3 -- determinant only (the inverse is unused here). */
static double invmtx_det(const double *a /* col-major */)
{
#define A(i, j) a[(i - 1) + 3 * (j - 1)]
    double t1 =  A(2,2) * A(3,3) - A(3,2) * A(2,3);
    double t2 = -A(2,1) * A(3,3) + A(3,1) * A(2,3);
    double t3 =  A(2,1) * A(3,2) - A(3,1) * A(2,2);
    return A(1,1) * t1 + A(1,2) * t2 + A(1,3) * t3;
#undef A
}

void orc_hex8_kinematics(const double *elcod, const double *eldis,
                         double *gpcar, double *gpvol, double *F, double *detF)
{
    double posgp[24], weigp[8], shapf[64], deriv[192];
    int ig, i, j, a;
    orc_hex8_tables(posgp, weigp, shapf, deriv);
    for (ig = 0; ig < 8; ig++) {
        double *gc = gpcar + 24 * ig, *Fg = F + 9 * ig;
        double gpdet = jacobi3(elcod, deriv + 24 * ig, gc);
        gpvol[ig] = weigp[ig] * gpdet;
        for (i = 0; i < 3; i++)
            for (j = 0; j < 3; j++) {
                double s = 0.0;
                for (a = 0; a < 8; a++) s = s + eldis[3 * a + i] * gc[3 * a + j];
                Fg[i + 3 * j] = s;
            }
        Fg[0] = Fg[0] + 1.0; Fg[4] = Fg[4] + 1.0; Fg[8] = Fg[8] + 1.0;
        detF[ig] = invmtx_det(Fg);
    }
}
