/*
 * TEST INFRASTRUCTURE ONLY -- CPU oracle of one explicit solidz step.
 *
 * Restates, FP64, in reference operation order:
 *   sld_updtss.f90:119-162 + elmlen.f90 (ndime==3) + sld_velsnd.f90:117-124
 *       critical time step  dt_cri = min_e hleng_min(e) / sqrt((lam+2mu)/rho)
 *   sld_funbou.f90:95-136 (itask 2)   DISCRETE[,SMOOTH] load tables
 *   sld_updbcs.f90:79-114             bvess <- f(t^{n+1}) on fixed dofs
 *   mod_sld_solution_methods.f90:101-223 (CD) and :385-574 (TW, prescribed
 *       displacement branch)          predictor / corrector
 *   sld_element_operations.f90:119-256 element loop: gather, elmca2,
 *       sld_elmpre, sld_elmcla -> sld_builtin_materials.f90:316-371 (sm160,
 *       P = F S), sld_elmmat.f90:115-209 (f_int, lumped mass)
 *   mod_matrix.f90:2327-2431          element -> nodal assembly
 *   solric.f90:294-330                a = r * (1/m)
 *   sld_updunk.f90:149-186            TIME_N <- ITER_K (history ping-pong)
 * The element loop runs in the order given by elem_order (colour-grouped when
 * color_ptr is set, with OpenMP across the elements of one colour, so each
 * node's contributions are summed in the same order as the GPU path).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

int orc_mutation = ORC_MUT_NONE;

#ifdef _OPENMP
#include <omp.h>
#endif
/* Threads of the colour-parallel element loop (the CPU baseline times 1 and
 * all of the host's cores). */
void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

static double elem_E(const orc_model *m, int64_t e, const double **props_out)
{
    int32_t mat = m->mat_of_elem ? m->mat_of_elem[e] : 0;
    *props_out = m->props + 18 * mat;
    return m->E_of_elem ? m->E_of_elem[e] : (*props_out)[0];
}

static void gather(const orc_model *m, const double *x, int64_t e, double *el)
{
    for (int a = 0; a < 8; a++) {
        int64_t p = m->lnods[8 * e + a];
        el[3 * a + 0] = x[3 * p + 0];
        el[3 * a + 1] = x[3 * p + 1];
        el[3 * a + 2] = x[3 * p + 2];
    }
}

int orc_sld_mass(const orc_model *m, orc_state *s)
{
    double posgp[24], weigp[8], shapf[64], deriv[192];
    orc_hex8_tables(posgp, weigp, shapf, deriv);
    memset(s->mass, 0, sizeof(double) * m->npoin);
    for (int64_t ee = 0; ee < m->nelem; ee++) {
        int64_t e = m->elem_order ? m->elem_order[ee] : ee;
        double elcod[24], zero[24] = {0}, gpcar[192], gpvol[8], F[72], detF[8], elmuu[8] = {0};
        const double *pr;
        int32_t mat = m->mat_of_elem ? m->mat_of_elem[e] : 0;
        (void)elem_E(m, e, &pr);
        gather(m, m->coord, e, elcod);
        orc_hex8_kinematics(elcod, zero, gpcar, gpvol, F, detF);
        /* sld_elmmat.f90:199-209 */
        for (int g = 0; g < 8; g++)
            for (int a = 0; a < 8; a++)
                for (int b = 0; b < 8; b++)
                    elmuu[a] = elmuu[a] + m->rho[mat] * gpvol[g] * shapf[8 * g + a] * shapf[8 * g + b];
        for (int a = 0; a < 8; a++) s->mass[m->lnods[8 * e + a]] += elmuu[a];
    }
    return 0;
}

double orc_sld_critical_dt(const orc_model *m)
{
    /* dercg: HEX08 derivatives at the centre (s=t=z=0) */
    double D[24];
    const double sgn[8][3] = {{-1,-1,-1},{1,-1,-1},{1,1,-1},{-1,1,-1},
                              {-1,-1,1},{1,-1,1},{1,1,1},{-1,1,1}};
    for (int a = 0; a < 8; a++)
        for (int d = 0; d < 3; d++) D[3 * a + d] = 0.125 * sgn[a][d];
    double dtmin = 1e300;
    for (int64_t e = 0; e < m->nelem; e++) {
        double elcod[24], xm[3][3] = {{0}}, tr[3][3], h[3];
        const double *pr;
        int32_t mat = m->mat_of_elem ? m->mat_of_elem[e] : 0;
        double E = elem_E(m, e, &pr), nu = pr[1];
        const int law = m->law_of_mat ? m->law_of_mat[mat] : 160;
        gather(m, m->coord, e, elcod);
        for (int k = 0; k < 8; k++)
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) xm[a][b] += elcod[3 * k + a] * D[3 * k + b];
        double t1 =  xm[1][1] * xm[2][2] - xm[2][1] * xm[1][2];
        double t2 = -xm[1][0] * xm[2][2] + xm[2][0] * xm[1][2];
        double t3 =  xm[1][0] * xm[2][1] - xm[2][0] * xm[1][1];
        double det = xm[0][0] * t1 + xm[0][1] * t2 + xm[0][2] * t3;
        double dn = 1.0 / det;
        tr[0][0] = t1 * dn; tr[1][0] = t2 * dn; tr[2][0] = t3 * dn;
        tr[1][1] = ( xm[0][0] * xm[2][2] - xm[2][0] * xm[0][2]) * dn;
        tr[2][1] = (-xm[0][0] * xm[2][1] + xm[0][1] * xm[2][0]) * dn;
        tr[2][2] = ( xm[0][0] * xm[1][1] - xm[1][0] * xm[0][1]) * dn;
        tr[0][1] = (-xm[0][1] * xm[2][2] + xm[2][1] * xm[0][2]) * dn;
        tr[0][2] = ( xm[0][1] * xm[1][2] - xm[1][1] * xm[0][2]) * dn;
        tr[1][2] = (-xm[0][0] * xm[1][2] + xm[1][0] * xm[0][2]) * dn;
        for (int i = 0; i < 3; i++)
            h[i] = 2.0 / sqrt(tr[i][0] * tr[i][0] + tr[i][1] * tr[i][1] + tr[i][2] * tr[i][2]);
        /* elmlen sorts hleng descending; hmini = hleng(ndime) / porde,
         * porde = 1 for HEX08 (sld_updtss.f90:143) */
        double hmin = fmin(h[0], fmin(h[1], h[2]));
        if (orc_mutation == ORC_MUT_DT_HMAX) hmin = fmax(h[0], fmax(h[1], h[2]));
        /* sld_velsnd.f90:117-124 (law 160): sqrt((lam + 2 mu)/rho) */
        double lam = nu * E / ((1.0 + nu) * (1.0 - 2.0 * nu));
        double mu = E / (2.0 * (1.0 + nu));
        double c = sqrt((lam + 2.0 * mu) / m->rho[mat]);
        if (orc_mutation == ORC_MUT_DT_YOUNG) c = sqrt(E / m->rho[mat]);
        if (law == 102) {  /* sld_velsnd.f90:75-78: sqrt((K0 + 4 mu0/3)/rho) */
            c = sqrt((pr[1] + 4.0 * pr[0] / 3.0) / m->rho[mat]);
            if (orc_mutation == ORC_MUT_DT_NH_BULK) c = sqrt(pr[1] / m->rho[mat]);
        }
        double dt = hmin / c;
        if (dt < dtmin) dtmin = dt;
    }
    return dtmin;
}

double orc_funbou(const orc_model *m, int32_t f, int c, double t)
{
    const double *tab = m->func_table + (size_t)f * m->func_maxpts * 4;
    int np = m->func_npts[f];
    double x = 0.0;
    for (int i = 0; i < np - 1; i++) {
        double t1 = tab[4 * i + 3], t2 = tab[4 * (i + 1) + 3];
        if (t >= t1 && t <= t2) {
            double p1 = tab[4 * i + c], p2 = tab[4 * (i + 1) + c], mm;
            if (m->func_smooth[f] == 1) {
                mm = (t - t1) / (t2 - t1);
                x = p1 + (p2 - p1) * (mm * mm * mm) * (10.0 - 15.0 * mm + 6.0 * (mm * mm));
            } else {
                mm = (p2 - p1) / (t2 - t1);
                x = p1 + mm * (t - t1);
            }
            break;
        }
    }
    double tlast = tab[4 * (np - 1) + 3];
    if (t > tlast) x = tab[4 * (np - 1) + c] * t / tlast;
    return x;
}

/* sld_funbou.f90:142-167 (itask 3, body forces): the row pair with
 * t1 <= t < t2, m = (p2 - p1)/(t2 - t1), b = p1 - m t1, value m t + b;
 * zero when t lies outside the table (the caller's fonew = 0). */
double orc_funbou_body(const orc_model *m, int32_t f, int c, double t)
{
    const double *tab = m->func_table + (size_t)f * m->func_maxpts * 4;
    int np = m->func_npts[f];
    for (int i = 0; i < np - 1; i++) {
        int upper = orc_mutation == ORC_MUT_BODY_LE ? t <= tab[4 * (i + 1) + 3]
                                                     : t < tab[4 * (i + 1) + 3];
        if (t >= tab[4 * i + 3] && upper) {
            double t1 = tab[4 * i + 3], t2 = tab[4 * (i + 1) + 3];
            double p1 = tab[4 * i + c], p2 = tab[4 * (i + 1) + c];
            double mm = (p2 - p1) / (t2 - t1);
            double b = p1 - mm * t1;
            return mm * t + b;
        }
    }
    return 0.0;
}

/* Boundary faces (sld_bouope.f90:88-277): per face, the geometry bocod of
 * its 4 nodes (reference, or current with follower loads), the value of the
 * step (sld_updbcs.f90:153-181: constant, or under transient conditions the
 * function's column x in component 1 and zero elsewhere, sld_funbou.f90
 * itask 1), then at each of the 2x2 Gauss points (mod_quadrature_gauss.f90:
 * 2361-2370, s outer) the tangents t_s, t_t of bouder (mod_bouder.f90:74-89),
 * n = t_s x t_t, eucta = |n|, the unit normal of chenor (chenor.f90, the
 * normalisation; its orientation test is unreachable), the traction (code 3:
 * value(1) n, code 6: value) and elfex += w eucta tract N_a on dofs that are
 * not prescribed.  QUA04 shape functions of mod_lagrange.f90 shape2. */
void orc_face_loads(const orc_model *m, const double *u, double t, double *ff)
{
    const double g = 0.577350269189626;
    const double gp[4][2] = {{-g, -g}, {-g, g}, {g, -g}, {g, g}};
    memset(ff, 0, sizeof(double) * 3 * m->npoin);
    for (int64_t f = 0; f < m->nface; f++) {
        const int32_t *nb = m->lnodb + 4 * f;
        double x[4][3], val[3], elfex[4][3] = {{0}};
        for (int a = 0; a < 4; a++)
            for (int d = 0; d < 3; d++) {
                x[a][d] = m->coord[3 * nb[a] + d];
                if (m->face_follower) x[a][d] = x[a][d] + u[3 * nb[a] + d];
            }
        int fn = m->face_func ? m->face_func[f] : -1;
        if (m->face_transient) {
            val[0] = fn >= 0 ? orc_funbou(m, fn, 0, t) : 0.0;
            val[1] = 0.0;
            val[2] = 0.0;
        } else {
            for (int d = 0; d < 3; d++) val[d] = m->face_value[3 * f + d];
        }
        for (int ig = 0; ig < 4; ig++) {
            double s = gp[ig][0], tt = gp[ig][1], st = s * tt;
            double N[4] = {(1.0 - tt - s + st) * 0.25, (1.0 - tt + s - st) * 0.25,
                           (1.0 + tt + s + st) * 0.25, (1.0 + tt - s - st) * 0.25};
            double ds[4] = {(-1.0 + tt) * 0.25, (1.0 - tt) * 0.25, (1.0 + tt) * 0.25,
                            (-1.0 - tt) * 0.25};
            double dt[4] = {(-1.0 + s) * 0.25, (-1.0 - s) * 0.25, (1.0 + s) * 0.25,
                            (1.0 - s) * 0.25};
            double t1[3], t2[3], n[3];
            for (int d = 0; d < 3; d++) {
                t1[d] = 0.0;
                t2[d] = 0.0;
                for (int a = 0; a < 4; a++) {
                    t1[d] = t1[d] + x[a][d] * ds[a];
                    t2[d] = t2[d] + x[a][d] * dt[a];
                }
            }
            n[0] = t1[1] * t2[2] - t1[2] * t2[1];
            n[1] = t1[2] * t2[0] - t1[0] * t2[2];
            n[2] = t1[0] * t2[1] - t1[1] * t2[0];
            double eucta = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            double gbsur = 1.0 * eucta, tract[3];
            if (m->face_code[f] == 3) {
                for (int d = 0; d < 3; d++) tract[d] = (n[d] / eucta) * val[0];
            } else {
                for (int d = 0; d < 3; d++) tract[d] = val[d];
            }
            for (int a = 0; a < 4; a++)
                for (int d = 0; d < 3; d++)
                    if (!((m->fixmask[nb[a]] >> d) & 1))
                        elfex[a][d] = elfex[a][d] + gbsur * tract[d] * N[a];
        }
        for (int a = 0; a < 4; a++)
            for (int d = 0; d < 3; d++) ff[3 * nb[a] + d] = ff[3 * nb[a] + d] + elfex[a][d];
    }
}

/* HEX08 shape function N_a at Gauss point g (mod_lagrange.f90 shape3) */
static double shapf_at(int g, int a)
{
    static double shapf[64];
    static int ready = 0;
    if (!ready) {
        double posgp[24], weigp[8], deriv[192];
        orc_hex8_tables(posgp, weigp, shapf, deriv);
        ready = 1;
    }
    return shapf[8 * g + a];
}

/* Element loop: fint, rhs and the new history/stress for every element.
 * vel = veloc_sld(:,:,ITER_K) at the time of the element loop (the Rayleigh
 * damping's velocity), NULL without damping. */
static void element_loop(const orc_model *m, orc_state *s, const double *vel)
{
    double *sdv_k = s->sdv, *sdv_n = s->sdv + 24 * m->nelem;
    int ncol = m->color_ptr ? m->ncolor : 1;
    (void)shapf_at(0, 0);   /* fill the table before the parallel region */
    memset(s->fint, 0, sizeof(double) * 3 * m->npoin);
    memset(s->rhs, 0, sizeof(double) * 3 * m->npoin);
    for (int c = 0; c < ncol; c++) {
        int64_t lo = m->color_ptr ? m->color_ptr[c] : 0;
        int64_t hi = m->color_ptr ? m->color_ptr[c + 1] : m->nelem;
        int64_t ee;
        #pragma omp parallel for schedule(dynamic, 64) if (m->color_ptr != NULL)
        for (ee = lo; ee < hi; ee++) {
            int64_t e = m->elem_order ? m->elem_order[ee] : ee;
            double elcod[24], eldis[24], gpcar[192], gpvol[8], F[72], detF[8];
            double elfint[24] = {0}, elrayl[24] = {0}, Sg[72];
            double props[18];
            const double *pr;
            double E = elem_E(m, e, &pr);
            const int32_t mat = m->mat_of_elem ? m->mat_of_elem[e] : 0;
            const int law = m->law_of_mat ? m->law_of_mat[mat] : 160;
            memcpy(props, pr, sizeof(props));
            if (law == 160) props[0] = E;
            ORC_PHASE(ORC_PH_KIN);
            gather(m, m->coord, e, elcod);
            gather(m, s->u, e, eldis);
            orc_hex8_kinematics(elcod, eldis, gpcar, gpvol, F, detF);
            /* Rayleigh damping (sld_elmmat.f90:160-163, :174-194) with the
             * velocity interpolated at the Gauss points, gpvel, and its
             * material gradient, gprat (sld_elmpre.f90:110-114, :150-157) */
            const double alpha = m->damping ? m->damping[2 * mat] : 0.0;
            const double beta = m->damping ? m->damping[2 * mat + 1] : 0.0;
            double gpvel[24], gprat[72];
            if (vel && (alpha > 0.0 || beta > 0.0)) {
                double elvel[24], posgp[24], weigp[8], shapf[64], deriv[192];
                orc_hex8_tables(posgp, weigp, shapf, deriv);
                gather(m, vel, e, elvel);
                for (int i = 0; i < 3; i++)
                    for (int g = 0; g < 8; g++) {
                        gpvel[3 * g + i] = 0.0;
                        for (int a = 0; a < 8; a++)
                            gpvel[3 * g + i] = gpvel[3 * g + i] + elvel[3 * a + i] * shapf[8 * g + a];
                    }
                for (int g = 0; g < 8; g++)
                    for (int i = 0; i < 3; i++)
                        for (int j = 0; j < 3; j++) {
                            double r = 0.0;
                            for (int a = 0; a < 8; a++)
                                r = r + elvel[3 * a + i] * gpcar[24 * g + 3 * a + j];
                            gprat[9 * g + i + 3 * j] = r;
                        }
            }
            for (int g = 0; g < 8; g++) {
                double S[9], P[9];
                const double *Fg = F + 9 * g;
                double *sk = sdv_k + 24 * e + 3 * g, *sn = sdv_n + 24 * e + 3 * g;
                uint32_t fl = 0;
                ORC_PHASE(ORC_PH_LAW);
                if (law == 102) {
                    /* no history: TIME_N carries over to ITER_K unchanged */
                    orc_nh102_point(props, Fg, detF[g], S);
                    for (int q = 0; q < 3; q++) sk[q] = sn[q];
                } else {
                    fl = orc_sm160_point(props, Fg, sn, sk, S, 0);
                }
                ORC_PHASE(ORC_PH_FORCE);
                if (!(detF[g] > 0.0)) fl |= ORC_FLAG_DETF_NONPOS;
                /* sld_elmpre.f90:229-240: the smallest det F <= 0 and its element */
                if (detF[g] <= 0.0) {
                    #pragma omp critical (orc_detf_min)
                    if (!s->halt || detF[g] < s->halt_detf ||
                        (detF[g] == s->halt_detf && e < s->halt_elem)) {
                        s->halt = 1;
                        s->halt_detf = detF[g];
                        s->halt_elem = e;
                    }
                }
                if (s->flags) s->flags[8 * e + g] = fl;
                if (s->S) memcpy(s->S + 72 * e + 9 * g, S, sizeof(S));
                memcpy(Sg + 9 * g, S, sizeof(S));
                /* P_jk = sum_i F_ji S_ik  (sld_builtin_materials.f90:359-369) */
                for (int q = 0; q < 9; q++) P[q] = 0.0;
                for (int k = 0; k < 3; k++)
                    for (int j = 0; j < 3; j++)
                        for (int i = 0; i < 3; i++)
                            P[j + 3 * k] = P[j + 3 * k] + Fg[j + 3 * i] * S[i + 3 * k];
                /* sld_elmmat.f90:137-150 */
                double volux = gpvol[g];
                for (int a = 0; a < 8; a++)
                    for (int i = 0; i < 3; i++) {
                        for (int j = 0; j < 3; j++)
                            elfint[3 * a + i] = elfint[3 * a + i]
                                + (P[i + 3 * j] + 0.0) * gpcar[24 * g + 3 * a + j] * volux;
                        /* sld_elmmat.f90:160-163 */
                        if (vel && alpha > 0.0) {
                            double vg = gpvel[3 * g + i];
                            if (orc_mutation == ORC_MUT_DAMP_LUMPED)
                                vg = vel[3 * m->lnods[8 * e + a] + i];
                            elrayl[3 * a + i] = elrayl[3 * a + i]
                                + alpha * vg * shapf_at(g, a) * m->rho[mat] * volux;
                        }
                    }
            }
            /* elrhs = elfext - elfint - elrayl_alpha (sld_elmmat.f90:170;
             * the element loads of gravity and the body force are the
             * lumped mass times the acceleration, added per node) */
            double elrhs[24];
            for (int q = 0; q < 24; q++) elrhs[q] = 0.0 - elfint[q] - elrayl[q];
            /* the beta term as written (sld_elmmat.f90:174-194): the same
             * sum over (j, k) is added to every component of node a */
            if (vel && beta > 0.0)
                for (int g = 0; g < 8; g++) {
                    double volux = gpvol[g];
                    for (int a = 0; a < 8; a++)
                        for (int i = 0; i < 3; i++)
                            for (int j = 0; j < 3; j++)
                                for (int k = 0; k < 3; k++)
                                    elrhs[3 * a + i] = elrhs[3 * a + i]
                                        + beta * gpcar[24 * g + 3 * a + j]
                                          * gprat[orc_mutation == ORC_MUT_BETA_T ? 9 * g + j + 3 * k
                                                                                 : 9 * g + k + 3 * j]
                                          * Sg[9 * g + j + 3 * k] * volux;
                }
            for (int a = 0; a < 8; a++) {
                int64_t p = m->lnods[8 * e + a];
                for (int i = 0; i < 3; i++) {
                    s->rhs[3 * p + i] += elrhs[3 * a + i];
                    s->fint[3 * p + i] += elfint[3 * a + i];
                }
            }
            ORC_PHASE(ORC_PH_NODAL);
        }
    }
}

int orc_sld_step(const orc_model *m, orc_state *s, double dt, int scheme, double phi)
{
    if (s->halt) return s->halt;   /* a stopped run stays stopped */
    int64_t np = m->npoin;
    double cutim = s->t + dt;             /* setgts.f90:98 */
    double t_end = cutim + dt, t_mid = 0.5 * (cutim + t_end);
    double fx = 0.0;
    if (m->fext0) fx = orc_funbou(m, m->fext_func, 0, cutim);
    double *u_old = NULL, *v_damp = NULL;
    if (m->damping) {
        v_damp = (double *)malloc(sizeof(double) * 3 * np);
        memcpy(v_damp, s->v, sizeof(double) * 3 * np);
    }
    if (s->fint_prev) {
        u_old = (double *)malloc(sizeof(double) * 3 * np);
        memcpy(u_old, s->u, sizeof(double) * 3 * np);
    }

    /* predictor */
    for (int64_t p = 0; p < np; p++) {
        for (int d = 0; d < 3; d++) {
            int fixed = (m->fixmask[p] >> d) & 1;
            double bv = 0.0;
            if (fixed && m->func_of_node && m->func_of_node[p] >= 0)
                bv = orc_funbou(m, m->func_of_node[p], d, cutim);
            double *u = &s->u[3 * p + d], *v = &s->v[3 * p + d], *a = &s->a[3 * p + d];
            if (scheme == 0) {
                double vh = *v + (t_mid - cutim) * *a;
                double un = *u + dt * vh;
                if (fixed) un = bv;
                *v = vh;
                *u = un;
            } else {
                /* mod_sld_solution_methods.f90:411-421 */
                double vn = *v + dt * *a;
                double un = *u + dt * *v + phi * (dt * dt) * *a;
                if (orc_mutation == ORC_MUT_TW_HALF) un = *u + dt * *v + 0.5 * phi * (dt * dt) * *a;
                if (orc_mutation == ORC_MUT_TW_VNEW) un = *u + dt * vn + phi * (dt * dt) * *a;
                if (fixed) un = bv;
                *v = vn;
                *u = un;
            }
        }
    }
    /* a non-finite predicted displacement stops the run (sld_cvgunk.f90:
     * 195-199, the displacement residual) */
    for (int64_t p = 0; p < np; p++)
        for (int d = 0; d < 3; d++)
            if (!isfinite(s->u[3 * p + d]) && !s->halt) {
                s->halt = 2;
                s->halt_node = p;
            }
    if (s->halt) {
        free(u_old);
        free(v_damp);
        return s->halt;
    }
    /* veloc_sld(ITER_K) at the element loop: v^n in the central difference
     * (its predictor keeps v^{n+1/2} in veloctmp_sld), the predicted
     * v^{n+1} in Tchamwa-Wielgosz */
    element_loop(m, s, m->damping ? ((scheme == 0 && orc_mutation != ORC_MUT_DAMP_VHALF)
                                     ? v_damp : s->v) : NULL);
    free(v_damp);
    v_damp = NULL;
    /* det F <= 0 stops the run after the element loop (sld_elmope.f90:106) */
    if (s->halt) {
        free(u_old);
        return s->halt;
    }
    /* f_ext = face loads (sld_bouope, assembled after the element loop in
     * sld_matrix.f90:221) + the nodal force pattern */
    double *fe = (double *)calloc((size_t)3 * np, sizeof(double));
    if (m->nface > 0) orc_face_loads(m, s->u, cutim, fe);
    if (m->fext0)
        for (int64_t p = 0; p < np; p++)
            for (int d = 0; d < 3; d++) fe[3 * p + d] = fe[3 * p + d] + fx * m->fext0[3 * p + d];
    /* gravity and the body force: (grnor gravi + bodyf) times the lumped
     * mass, on every dof, prescribed ones too (their reactions) */
    if (m->has_gravity || m->body_func >= 0) {
        double gb[3];
        for (int d = 0; d < 3; d++) {
            double g = m->has_gravity ? m->gravity[d] : 0.0;
            double b = m->body_func >= 0 ? orc_funbou_body(m, m->body_func, d, cutim) : 0.0;
            gb[d] = g + b;
        }
        for (int64_t p = 0; p < np; p++)
            for (int d = 0; d < 3; d++) fe[3 * p + d] = fe[3 * p + d] + s->mass[p] * gb[d];
    }
    for (int64_t p = 0; p < 3 * np; p++) s->rhs[p] = s->rhs[p] + fe[p];
    /* a non-finite residual stops the run after the step (sld_cvgunk.f90:195-199) */
    int64_t bad_node = -1;
    for (int64_t p = np - 1; p >= 0; p--)
        for (int d = 0; d < 3; d++)
            if (!isfinite(s->rhs[3 * p + d])) bad_node = p;
    /* solve + corrector */
    for (int64_t p = 0; p < np; p++) {
        double fact1 = 1.0 / s->mass[p];
        for (int d = 0; d < 3; d++) {
            double acc = 0.0 + s->rhs[3 * p + d] * fact1;
            int fixed = (m->fixmask[p] >> d) & 1;
            if (scheme == 0) {
                s->a[3 * p + d] = acc;
                s->v[3 * p + d] = s->v[3 * p + d] + (t_end - t_mid) * acc;
            } else {
                /* mod_sld_solution_methods.f90:504-514 */
                s->a[3 * p + d] = acc;
                if (fixed && orc_mutation != ORC_MUT_TW_FIXVA) {
                    s->v[3 * p + d] = 0.0;
                    s->a[3 * p + d] = 0.0;
                }
            }
        }
    }
    /* energies (mod_sld_energy.f90:113-140): trapezoidal internal energy
     * and external work over the displacement increment (full weight on
     * the first step), kinetic energy of the lumped mass */
    if (s->fint_prev) {
        double ie = 0.0, wk = 0.0, ke = 0.0;
        for (int64_t p = 0; p < np; p++)
            for (int d = 0; d < 3; d++) {
                double du = s->u[3 * p + d] - u_old[3 * p + d];
                ie = ie + (s->fint[3 * p + d] + s->fint_prev[3 * p + d]) * du;
                wk = wk + (fe[3 * p + d] + s->fext_prev[3 * p + d]) * du;
                ke = ke + s->v[3 * p + d] * s->v[3 * p + d] * s->mass[p];
            }
        double w = s->nsteps == 0 ? 1.0 : 0.5;
        s->allie = s->allie + w * fabs(ie);
        s->allwk = s->allwk + w * fabs(wk);
        s->allke = 0.5 * ke;
        s->etota = s->allie + s->allke - s->allwk;
        memcpy(s->fint_prev, s->fint, sizeof(double) * 3 * np);
        memcpy(s->fext_prev, fe, sizeof(double) * 3 * np);
        free(u_old);
    }
    free(fe);
    /* sld_updunk ENDSTE: TIME_N <- ITER_K */
    memcpy(s->sdv + 24 * m->nelem, s->sdv, sizeof(double) * 24 * m->nelem);
    s->t = cutim;
    s->nsteps++;
    if (bad_node >= 0) {
        s->halt = 2;
        s->halt_node = bad_node;
    }
    return s->halt;
}
