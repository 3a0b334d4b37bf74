/*
 * efem_oracle.c -- CPU restatement of EFEM's Total-Lagrangian solve loop.
 *
 * TEST INFRASTRUCTURE ONLY (see efem_oracle.h).  Plain C99, FP64, one thread.
 * Reference paths are relative to the SeanPei/EFEM source tree.
 */
#define _POSIX_C_SOURCE 199309L
#include "efem_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ======================================================================== */
/* small 3x3 helpers (row-major)                                            */
/* ======================================================================== */
static double det3(const double* A)
{
    return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6])
        + A[2] * (A[3] * A[7] - A[4] * A[6]);
}

/* inverse through the adjugate, as a cofactor expansion divided by det */
static double inv3(const double* A, double* B)
{
    double c00 = A[4] * A[8] - A[5] * A[7];
    double c01 = A[5] * A[6] - A[3] * A[8];
    double c02 = A[3] * A[7] - A[4] * A[6];
    double d = A[0] * c00 + A[1] * c01 + A[2] * c02;
    double id = 1.0 / d;
    B[0] = c00 * id;
    B[1] = (A[2] * A[7] - A[1] * A[8]) * id;
    B[2] = (A[1] * A[5] - A[2] * A[4]) * id;
    B[3] = c01 * id;
    B[4] = (A[0] * A[8] - A[2] * A[6]) * id;
    B[5] = (A[2] * A[3] - A[0] * A[5]) * id;
    B[6] = c02 * id;
    B[7] = (A[1] * A[6] - A[0] * A[7]) * id;
    B[8] = (A[0] * A[4] - A[1] * A[3]) * id;
    return d;
}

static void cross3(const double* a, const double* b, double* c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

static double now_s(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* ======================================================================== */
/* element library                                                          */
/* ======================================================================== */

/* C3D8 nodal parametric coordinates: VTK hexahedron order
 * (HexahedraElement.cpp:15-58 uses these signs node by node). */
static const double HEX8_NODE[8][3] = {
    { -1, -1, -1 }, { 1, -1, -1 }, { 1, 1, -1 }, { -1, 1, -1 },
    { -1, -1, 1 }, { 1, -1, 1 }, { 1, 1, 1 }, { -1, 1, 1 }
};

/* C3D20 nodal parametric coordinates in the reference's own parametrisation
 * (HexahedraElement.cpp:147-238): node 0 sits at (1,-1,1); in VTK (r,s,t)
 * terms (xi,eta,zeta) = (-t, r, -s). */
static const double HEX20_NODE[20][3] = {
    { 1, -1, 1 }, { 1, 1, 1 }, { 1, 1, -1 }, { 1, -1, -1 },
    { -1, -1, 1 }, { -1, 1, 1 }, { -1, 1, -1 }, { -1, -1, -1 },
    { 1, 0, 1 }, { 1, 1, 0 }, { 1, 0, -1 }, { 1, -1, 0 },
    { -1, 0, 1 }, { -1, 1, 0 }, { -1, 0, -1 }, { -1, -1, 0 },
    { 0, -1, 1 }, { 0, 1, 1 }, { 0, 1, -1 }, { 0, -1, -1 }
};

/* C3D10 edge nodes: barycentric pairs (TetrahedraElement.cpp:62-79);
 * barycentric l0 = 1-xi-eta-zeta (node 0), l1 = xi, l2 = eta, l3 = zeta */
static const int TET10_EDGE[6][2] = { { 0, 1 }, { 1, 2 }, { 2, 0 }, { 0, 3 }, { 1, 3 }, { 2, 3 } };

/* N (nen) and dN/dxi ([nen][3]) at xi.  Returns 0, or -1 for an unknown nen.
 * C3D4: N = (xi, eta, zeta, 1-xi-eta-zeta)  TetrahedraElement.cpp:15-33
 * C3D10: quadratic Lagrange tet             TetrahedraElement.cpp:62-122
 * C3D8: trilinear                           HexahedraElement.cpp:15-58
 * C3D20: serendipity                        HexahedraElement.cpp:147-238 */
int or_shape(int nen, const double xi[3], double* N, double* dN)
{
    if (nen == 4) {
        N[0] = xi[0];
        N[1] = xi[1];
        N[2] = xi[2];
        N[3] = 1.0 - xi[0] - xi[1] - xi[2];
        static const double g[4][3] = { { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 }, { -1, -1, -1 } };
        memcpy(dN, g, sizeof(g));
        return 0;
    }
    if (nen == 10) {
        double l[4] = { 1.0 - xi[0] - xi[1] - xi[2], xi[0], xi[1], xi[2] };
        static const double dl[4][3] = { { -1, -1, -1 }, { 1, 0, 0 }, { 0, 1, 0 }, { 0, 0, 1 } };
        for (int i = 0; i < 4; i++) {
            N[i] = l[i] * (2.0 * l[i] - 1.0);
            for (int d = 0; d < 3; d++)
                dN[3 * i + d] = (4.0 * l[i] - 1.0) * dl[i][d];
        }
        for (int e = 0; e < 6; e++) {
            int a = TET10_EDGE[e][0], b = TET10_EDGE[e][1];
            N[4 + e] = 4.0 * l[a] * l[b];
            for (int d = 0; d < 3; d++)
                dN[3 * (4 + e) + d] = 4.0 * (l[a] * dl[b][d] + l[b] * dl[a][d]);
        }
        return 0;
    }
    if (nen == 8) {
        for (int a = 0; a < 8; a++) {
            double s0 = HEX8_NODE[a][0], s1 = HEX8_NODE[a][1], s2 = HEX8_NODE[a][2];
            double f0 = 1.0 + s0 * xi[0], f1 = 1.0 + s1 * xi[1], f2 = 1.0 + s2 * xi[2];
            N[a] = 0.125 * f0 * f1 * f2;
            dN[3 * a + 0] = 0.125 * s0 * f1 * f2;
            dN[3 * a + 1] = 0.125 * f0 * s1 * f2;
            dN[3 * a + 2] = 0.125 * f0 * f1 * s2;
        }
        return 0;
    }
    if (nen == 20) {
        for (int a = 0; a < 20; a++) {
            double s[3] = { HEX20_NODE[a][0], HEX20_NODE[a][1], HEX20_NODE[a][2] };
            if (a < 8) {
                double f0 = 1.0 + s[0] * xi[0], f1 = 1.0 + s[1] * xi[1], f2 = 1.0 + s[2] * xi[2];
                double g = s[0] * xi[0] + s[1] * xi[1] + s[2] * xi[2] - 2.0;
                N[a] = f0 * f1 * f2 * g / 8.0;
                dN[3 * a + 0] = s[0] * f1 * f2 * (g + f0) / 8.0;
                dN[3 * a + 1] = s[1] * f0 * f2 * (g + f1) / 8.0;
                dN[3 * a + 2] = s[2] * f0 * f1 * (g + f2) / 8.0;
            } else {
                /* the zero coordinate k carries (1 - xi_k^2) */
                int k = (s[0] == 0.0) ? 0 : ((s[1] == 0.0) ? 1 : 2);
                int i = (k + 1) % 3, j = (k + 2) % 3;
                double q = 1.0 - xi[k] * xi[k];
                double fi = 1.0 + s[i] * xi[i], fj = 1.0 + s[j] * xi[j];
                N[a] = q * fi * fj / 4.0;
                dN[3 * a + k] = -2.0 * xi[k] * fi * fj / 4.0;
                dN[3 * a + i] = q * s[i] * fj / 4.0;
                dN[3 * a + j] = q * fi * s[j] / 4.0;
            }
        }
        return 0;
    }
    return -1;
}

/* Gauss rules, IntegralElement.inl:3-223.  Weights are normalised to sum 1
 * (they multiply the element VOLUME, ElementConstraint.cpp:148).
 * Full rules (the only ones the reference instantiates, Constraint.cpp:9-20):
 *   C3D4 1 pt, C3D10 4 pt, C3D8 2x2x2, C3D20 3x3x3.
 * Reduced rules (the reference's unreachable *R classes):
 *   C3D4R 1 pt, C3D10R 1 pt, C3D8R 1 pt, C3D20R 2x2x2. */
int or_n_gauss(int nen, int reduced)
{
    switch (nen) {
    case 4: return 1;
    case 10: return reduced ? 1 : 4;
    case 8: return reduced ? 1 : 8;
    case 20: return reduced ? 8 : 27;
    }
    return 0;
}

static void tet_rule(int n, int q, double xi[3], double* w)
{
    if (n == 1) {
        xi[0] = xi[1] = xi[2] = 0.25;
        *w = 1.0;
        return;
    }
    /* truncated 8-digit constants of IntegralElement.inl:41-47 */
    const double a = 0.58541020, b = 0.13819660;
    xi[0] = xi[1] = xi[2] = b;
    if (q < 3)
        xi[q] = a;
    *w = 0.25;
}

static void hex_rule(int n, int q, double xi[3], double* w)
{
    if (n == 1) {
        xi[0] = xi[1] = xi[2] = 0.0;
        *w = 1.0;
        return;
    }
    if (n == 8) {
        /* sign pattern of IntegralElement.inl:148-164 */
        static const int sg[8][3] = { { -1, -1, -1 }, { 1, -1, -1 }, { 1, 1, -1 }, { -1, 1, -1 },
            { -1, -1, 1 }, { 1, -1, 1 }, { 1, 1, 1 }, { -1, 1, 1 } };
        double g = 1.00 / sqrt(3.0);
        for (int d = 0; d < 3; d++)
            xi[d] = sg[q][d] * g;
        *w = 0.125;
        return;
    }
    /* 27 points, xi fastest (IntegralElement.inl:165-220), weights
     * {5,8,5}^3 / 5832 (IntegralElement.inl:65-123) */
    double r = sqrt(3.00 / 5.00);
    const double pt[3] = { -r, 0.0, r };
    const int wt[3] = { 5, 8, 5 };
    int i = q % 3, j = (q / 3) % 3, k = q / 9;
    xi[0] = pt[i];
    xi[1] = pt[j];
    xi[2] = pt[k];
    *w = (double)(wt[i] * wt[j] * wt[k]) / 5832.0;
}

int or_gauss(int nen, int reduced, int q, double xi[3], double* w)
{
    int n = or_n_gauss(nen, reduced);
    if (q < 0 || q >= n)
        return -1;
    if (nen == 4 || nen == 10)
        tet_rule(n, q, xi, w);
    else
        hex_rule(n, q, xi, w);
    return n;
}

/* ------------------------------------------------------------------------ */
/* element volume: Element::volume (model/Element.h:25-89) takes the VTK     */
/* surface of the one cell, triangulates it and returns                      */
/* vtkMassProperties::GetVolumeProjected.  For a closed, consistently        */
/* oriented triangulated surface this is the divergence-theorem volume.      */
/* Quadratic faces are tessellated through their mid-edge nodes.             */
/* ------------------------------------------------------------------------ */
static const int TET_FACES[4][3] = { { 0, 1, 3 }, { 1, 2, 3 }, { 2, 0, 3 }, { 0, 2, 1 } };
static const int HEX_FACES[6][4] = { { 0, 4, 7, 3 }, { 1, 2, 6, 5 }, { 0, 1, 5, 4 },
    { 3, 7, 6, 2 }, { 0, 3, 2, 1 }, { 4, 5, 6, 7 } };
static const int HEX20_EDGE[12][2] = { { 0, 1 }, { 1, 2 }, { 2, 3 }, { 3, 0 }, { 4, 5 },
    { 5, 6 }, { 6, 7 }, { 7, 4 }, { 0, 4 }, { 1, 5 }, { 2, 6 }, { 3, 7 } };

static int mid_node(int nen, int a, int b)
{
    if (nen == 10) {
        for (int e = 0; e < 6; e++)
            if ((TET10_EDGE[e][0] == a && TET10_EDGE[e][1] == b) || (TET10_EDGE[e][0] == b && TET10_EDGE[e][1] == a))
                return 4 + e;
    } else if (nen == 20) {
        for (int e = 0; e < 12; e++)
            if ((HEX20_EDGE[e][0] == a && HEX20_EDGE[e][1] == b) || (HEX20_EDGE[e][0] == b && HEX20_EDGE[e][1] == a))
                return 8 + e;
    }
    return -1;
}

static double tri_vol(const double* Xe, int a, int b, int c)
{
    double bc[3];
    cross3(Xe + 3 * b, Xe + 3 * c, bc);
    const double* pa = Xe + 3 * a;
    return pa[0] * bc[0] + pa[1] * bc[1] + pa[2] * bc[2];
}

double or_element_volume(int nen, const double* Xe)
{
    double s = 0.0;
    if (nen == 4 || nen == 10) {
        for (int f = 0; f < 4; f++) {
            int a = TET_FACES[f][0], b = TET_FACES[f][1], c = TET_FACES[f][2];
            if (nen == 4) {
                s += tri_vol(Xe, a, b, c);
            } else {
                int ab = mid_node(nen, a, b), bc = mid_node(nen, b, c), ca = mid_node(nen, c, a);
                s += tri_vol(Xe, a, ab, ca) + tri_vol(Xe, ab, b, bc) + tri_vol(Xe, ca, bc, c)
                    + tri_vol(Xe, ab, bc, ca);
            }
        }
    } else {
        for (int f = 0; f < 6; f++) {
            int a = HEX_FACES[f][0], b = HEX_FACES[f][1], c = HEX_FACES[f][2], d = HEX_FACES[f][3];
            if (nen == 8) {
                s += tri_vol(Xe, a, b, c) + tri_vol(Xe, a, c, d);
            } else {
                int ab = mid_node(nen, a, b), bc = mid_node(nen, b, c), cd = mid_node(nen, c, d),
                    da = mid_node(nen, d, a);
                s += tri_vol(Xe, a, ab, da) + tri_vol(Xe, b, bc, ab) + tri_vol(Xe, c, cd, bc)
                    + tri_vol(Xe, d, da, cd) + tri_vol(Xe, ab, bc, cd) + tri_vol(Xe, ab, cd, da);
            }
        }
    }
    return fabs(s) / 6.0;
}

/* ======================================================================== */
/* materials                                                                */
/* ======================================================================== */

/* PK1 stress.  NeoHookean.cpp:4-20, MooneyRivlin.cpp:4-24, Yeoh.cpp:4-20:
 *   I1bar = I1 / I3^(1/3),  dJ/dF = J F^-T,
 *   P_I1bar = -2 I1bar dJdF / (3J) + 2F / I3^(1/3)
 *   P_I2bar = -4 I2bar dJdF / (3J) + 2F(I1 I - C) / (I3^2)^(1/3)
 *   P = c1 P_I1bar [+ C01 P_I2bar] + (2/D1)(J-1) dJdF
 * with c1 = C10 (NH, MR) or C10 + 2 C20 (I1bar-3) + 3 C30 (I1bar-3)^2 (Yeoh). */
void or_piola(int law, const double p[4], const double F[9], double P[9])
{
    double J = det3(F);
    double C[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            C[3 * i + j] = F[i] * F[j] + F[3 + i] * F[3 + j] + F[6 + i] * F[6 + j];
    double I1 = C[0] + C[4] + C[8];
    double I3 = J * J;
    double c13 = pow(I3, 1.0 / 3.0);
    double I1bar = I1 / c13;
    double Fi[9];
    inv3(F, Fi);
    double dJdF[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            dJdF[3 * i + j] = J * Fi[3 * j + i];
    double c1 = p[0];
    if (law == OR_YEOH) {
        double t = I1bar - 3.0;
        c1 = p[0] + 2.0 * p[1] * t + 3.0 * p[2] * pow(t, 2);
    }
    for (int k = 0; k < 9; k++)
        P[k] = c1 * (-2.0 * I1bar * dJdF[k] / (3.0 * J) + 2.0 * F[k] / c13);
    if (law == OR_MOONEY_RIVLIN) {
        double trCC = 0.0;
        for (int k = 0; k < 9; k++)
            trCC += C[k] * C[k];
        double I2 = 0.5 * (pow(I1, 2) - trCC);
        double c23 = pow(I3 * I3, 1.0 / 3.0);
        double I2bar = I2 / c23;
        double M[9]; /* I1 I - C */
        for (int k = 0; k < 9; k++)
            M[k] = -C[k];
        M[0] += I1;
        M[4] += I1;
        M[8] += I1;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double dI2 = 2.0 * (F[3 * i] * M[j] + F[3 * i + 1] * M[3 + j] + F[3 * i + 2] * M[6 + j]);
                P[3 * i + j] += p[1] * (-4.0 * I2bar * dJdF[3 * i + j] / (3.0 * J) + dI2 / c23);
            }
    }
    double kv = 2.0 / p[3] * (J - 1.0);
    for (int k = 0; k < 9; k++)
        P[k] += kv * dJdF[k];
}

/* Strain energy density of the same invariants (an addition: the
 * reference has no energy).  W = c1-part + (J-1)^2/D1 with
 *   NH:   C10 (I1bar - 3)
 *   MR:   C10 (I1bar - 3) + C01 (I2bar - 3)
 *   Yeoh: C10 t + C20 t^2 + C30 t^3, t = I1bar - 3,
 * I1bar = I1 / (J^2)^(1/3), I2bar = I2 / (J^2)^(2/3) as in or_piola, so that
 * dW/dF = or_piola (pinned by central differences in the tests). */
double or_energy_density(int law, const double p[4], const double F[9])
{
    double J = det3(F);
    double C[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            C[3 * i + j] = F[i] * F[j] + F[3 + i] * F[3 + j] + F[6 + i] * F[6 + j];
    double I1 = C[0] + C[4] + C[8];
    double I3 = J * J;
    double t = I1 / pow(I3, 1.0 / 3.0) - 3.0;
    double W = (J - 1.0) * (J - 1.0) / p[3];
    if (law == OR_YEOH)
        return W + p[0] * t + p[1] * t * t + p[2] * t * t * t;
    W += p[0] * t;
    if (law == OR_MOONEY_RIVLIN) {
        double trCC = 0.0;
        for (int k = 0; k < 9; k++)
            trCC += C[k] * C[k];
        double I2 = 0.5 * (I1 * I1 - trCC);
        W += p[1] * (I2 / pow(I3 * I3, 1.0 / 3.0) - 3.0);
    }
    return W;
}

/* calibrated approximate-Jacobian coefficient, the "t=2.0" lines of
 * NeoHookean.cpp:41, MooneyRivlin.cpp:48, Yeoh.cpp:47 */
double or_elasticity_coeff(int law, const double p[4])
{
    if (law == OR_NEOHOOKEAN)
        return 2.94001 * p[0] + 2.0 / p[3];
    if (law == OR_MOONEY_RIVLIN)
        return 3.818 * p[1] + 2.94001 * p[0] + 2.0 / p[3];
    return 2.94001 * p[0] + 3.51194 * p[1] + 3.3154 * p[2] + 2.0 / p[3];
}

/* ======================================================================== */
/* model                                                                    */
/* ======================================================================== */
typedef struct {
    int elem;       /* parent element */
    int nq;
    double A0[3];
    double w[9];
    double N[9][20];
    double dN[9][20][3];
    double J0inv[9][9];
} or_face;

typedef struct {
    int n;          /* order of the factored matrix */
    int* perm;      /* perm[new] = old */
    int* Lp;
    int* Li;
    double* Lx;
    double* work;
} or_chol;

struct or_model {
    int n_nodes, n_elems, nen, ng, law, ndof, n_free;
    double* X;
    int* conn;
    int* fid;         /* free id per node or -1 */
    double* params;   /* [n_elems][4] */
    double* coeff;    /* [n_elems] elasticity_coeff */
    double* vol;      /* [n_elems] */
    double* gw;       /* [ng] */
    double* gdN;      /* [ng][nen][3] */
    double* J0inv;    /* [n_elems][ng][9] */
    double* T;        /* [n_elems][ng][nen][3] : grad_X N_a */
    int n_faces, fnn;
    or_face* faces;
    double pressure;
    int directed;
    double pdir[3];
    int n_fnodes;
    int* fnodes;
    double fvec[3];
    double* grav;     /* [ndof] gravity residual -M g, or NULL */
    or_chol* chol;
    int chol_scalar;
};

static void set_err(char* err, int len, const char* msg)
{
    if (err && len > 0) {
        strncpy(err, msg, (size_t)len - 1);
        err[len - 1] = 0;
    }
}

/* Gauss points of a pressure face on its parent element,
 * ExternalPressureFace::buildGaussPoint (ExternalConstraint.cpp:289-827). */
static int build_face(or_model* m, const int* fn, int fnn, or_face* F, char* err, int errlen)
{
    /* parent: the unique element holding every face node (:255-279) */
    int parent = -1, count = 0;
    for (int e = 0; e < m->n_elems; e++) {
        const int* c = m->conn + (size_t)e * m->nen;
        int all = 1;
        for (int i = 0; i < fnn && all; i++) {
            int found = 0;
            for (int a = 0; a < m->nen; a++)
                if (c[a] == fn[i]) {
                    found = 1;
                    break;
                }
            all = found;
        }
        if (all) {
            parent = e;
            count++;
        }
    }
    if (count != 1) {
        set_err(err, errlen, "pressure face: no unique parent element");
        return -1;
    }
    F->elem = parent;
    const int* c = m->conn + (size_t)parent * m->nen;
    int absent[20];
    for (int a = 0; a < m->nen; a++) {
        absent[a] = 1;
        for (int i = 0; i < fnn; i++)
            if (fn[i] == c[a])
                absent[a] = 0;
    }
    const double* X = m->X;
#define XN(k) (X + 3 * (size_t)(k))
    double v1[3], v2[3];
    double xi[9][3];
    int nq = 0;
    if (fnn == 3 || fnn == 6) {
        if ((fnn == 3 && m->nen != 4) || (fnn == 6 && m->nen != 10)) {
            set_err(err, errlen, "pressure face: node count does not match the parent element");
            return -1;
        }
        int k = -1;
        for (int a = 0; a < 4; a++)
            if (absent[a]) {
                k = a;
                break;
            }
        if (k < 0) {
            set_err(err, errlen, "pressure face: degenerate");
            return -1;
        }
        /* A0 = 1/2 (x2 - x0) x (x1 - x0), flipped toward the absent vertex */
        for (int d = 0; d < 3; d++) {
            v1[d] = XN(fn[2])[d] - XN(fn[0])[d];
            v2[d] = XN(fn[1])[d] - XN(fn[0])[d];
        }
        cross3(v1, v2, F->A0);
        for (int d = 0; d < 3; d++)
            F->A0[d] *= 0.5;
        double t = 0;
        for (int d = 0; d < 3; d++)
            t += (XN(c[k])[d] - XN(fn[0])[d]) * F->A0[d];
        if (t < 0)
            for (int d = 0; d < 3; d++)
                F->A0[d] = -F->A0[d];
        if (fnn == 3) {
            /* one point at the face centroid (:292-337) */
            nq = 1;
            F->w[0] = 1;
            for (int d = 0; d < 3; d++)
                xi[0][d] = 1.0 / 3.0;
            if (k < 3)
                xi[0][k] = 0;
        } else {
            /* three points (2/3,1/6,1/6) on the face (:338-409) */
            static const double T0[3][3] = { { 2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0 },
                { 1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0 }, { 1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0 } };
            static const double T1[3][3] = { { 0, 2.0 / 3.0, 1.0 / 6.0 }, { 0, 1.0 / 6.0, 2.0 / 3.0 },
                { 0, 1.0 / 6.0, 1.0 / 6.0 } };
            static const double T2[3][3] = { { 2.0 / 3.0, 0, 1.0 / 6.0 }, { 1.0 / 6.0, 0, 2.0 / 3.0 },
                { 1.0 / 6.0, 0, 1.0 / 6.0 } };
            static const double T3[3][3] = { { 2.0 / 3.0, 1.0 / 6.0, 0 }, { 1.0 / 6.0, 2.0 / 3.0, 0 },
                { 1.0 / 6.0, 1.0 / 6.0, 0 } };
            const double(*tab)[3] = k == 0 ? T0 : (k == 1 ? T1 : (k == 2 ? T2 : T3));
            nq = 3;
            for (int q = 0; q < 3; q++) {
                F->w[q] = 1.0 / 3.0;
                for (int d = 0; d < 3; d++)
                    xi[q][d] = tab[q][d];
            }
        }
    } else if (fnn == 4 || fnn == 8) {
        if ((fnn == 4 && m->nen != 8) || (fnn == 8 && m->nen != 20)) {
            set_err(err, errlen, "pressure face: node count does not match the parent element");
            return -1;
        }
        /* the face is identified by two element nodes missing from it
         * (:419-563 and :580-814); each case fixes one parametric axis */
        static const int Q4[6][5] = { { 0, 2, 2, 1, 0 }, { 0, 5, 1, 1, 0 }, { 0, 7, 0, 1, 0 },
            { 6, 3, 1, -1, 6 }, { 6, 1, 0, -1, 6 }, { 6, 4, 2, -1, 6 } };
        static const int Q8[6][5] = { { 0, 2, 0, -1, 0 }, { 0, 5, 2, -1, 0 }, { 0, 7, 1, 1, 0 },
            { 6, 3, 2, 1, 6 }, { 6, 1, 1, -1, 6 }, { 6, 4, 0, 1, 6 } };
        const int(*cs)[5] = fnn == 4 ? Q4 : Q8;
        int ci = -1;
        for (int i = 0; i < 6; i++)
            if (absent[cs[i][0]] && absent[cs[i][1]]) {
                ci = i;
                break;
            }
        if (ci < 0) {
            set_err(err, errlen, "pressure face: not a face of its parent element");
            return -1;
        }
        int axis = cs[ci][2], flip = cs[ci][4];
        double val = cs[ci][3];
        int f0 = (axis + 1) % 3 < (axis + 2) % 3 ? (axis + 1) % 3 : (axis + 2) % 3;
        int f1 = 3 - axis - f0;
        if (fnn == 4) {
            /* 2x2 points, first free axis (-,+,+,-), second (-,-,+,+) */
            double s = sqrt(1.0 / 3.0);
            static const int a0[4] = { -1, 1, 1, -1 }, a1[4] = { -1, -1, 1, 1 };
            nq = 4;
            for (int q = 0; q < 4; q++) {
                xi[q][axis] = val;
                xi[q][f0] = a0[q] * s;
                xi[q][f1] = a1[q] * s;
                F->w[q] = 0.25;
            }
        } else {
            /* 3x3 points, first free axis slowest; weights {25,40,64}/81/4 */
            double r = sqrt(0.6);
            const double pt[3] = { -r, 0.0, r };
            static const double wf[9] = { 25, 40, 25, 40, 64, 40, 25, 40, 25 };
            nq = 9;
            for (int q = 0; q < 9; q++) {
                xi[q][axis] = val;
                xi[q][f0] = pt[q / 3];
                xi[q][f1] = pt[q % 3];
                F->w[q] = 0.25 * (wf[q] / 81.0);
            }
        }
        /* A0 = A01 + A02 over the first four face nodes, each half flipped
         * toward an element node that is not on the face */
        double A01[3], A02[3];
        for (int d = 0; d < 3; d++) {
            v1[d] = XN(fn[2])[d] - XN(fn[0])[d];
            v2[d] = XN(fn[1])[d] - XN(fn[0])[d];
        }
        cross3(v1, v2, A01);
        for (int d = 0; d < 3; d++) {
            v1[d] = XN(fn[0])[d] - XN(fn[3])[d];
            v2[d] = XN(fn[2])[d] - XN(fn[3])[d];
        }
        cross3(v1, v2, A02);
        double t1 = 0, t2 = 0;
        for (int d = 0; d < 3; d++) {
            A01[d] *= 0.5;
            A02[d] *= 0.5;
        }
        for (int d = 0; d < 3; d++) {
            double dd = XN(c[flip])[d] - XN(fn[0])[d];
            t1 += dd * A01[d];
            t2 += dd * A02[d];
        }
        for (int d = 0; d < 3; d++)
            F->A0[d] = (t1 < 0 ? -A01[d] : A01[d]) + (t2 < 0 ? -A02[d] : A02[d]);
    } else {
        set_err(err, errlen, "pressure face: unsupported node count");
        return -1;
    }
    F->nq = nq;
    for (int q = 0; q < nq; q++) {
        double N[20], dN[60], J0[9];
        or_shape(m->nen, xi[q], N, dN);
        memset(J0, 0, sizeof(J0));
        for (int a = 0; a < m->nen; a++)
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++)
                    J0[3 * i + j] += XN(c[a])[i] * dN[3 * a + j];
        inv3(J0, F->J0inv[q]);
        for (int a = 0; a < m->nen; a++) {
            F->N[q][a] = N[a];
            for (int d = 0; d < 3; d++)
                F->dN[q][a][d] = dN[3 * a + d];
        }
    }
#undef XN
    return 0;
}

/* Constraint::Constraint (Constraint.cpp:22-91), IntegralElement::initial
 * (IntegralElement.cpp:3-19), ElementConstraint::ElementConstraint
 * (ElementConstraint.cpp:69-83) */
/* Gravity as the reference assembles it: per element a consistent mass
 * matrix M_e = V sum_q w_q rho N^T N (3x3 identity blocks) and a lumped one
 * (max(V rho, 4e-12) / nen on the diagonal), ElementConstraint::conInitial
 * (ElementConstraint.cpp:85-137), kept for free-free pairs only; the
 * gravity vector g repeated on every free node (InertiaConstraint::
 * buildGravity, InertiaConstraint.cpp:60-84); residual -M g
 * (conEvaluteResidual, :34-49). */
static void build_gravity(or_model* m, const or_problem* pb)
{
    const int nen = m->nen, ng = m->ng;
    double g[3] = { 0, 0, 0 };
    g[(pb->gravity - 1) / 2] = (pb->gravity % 2 == 1) ? 9.8 : -9.8;
    m->grav = (double*)calloc((size_t)m->ndof, sizeof(double));
    double Ng[27][20];
    for (int q = 0; q < ng; q++) {
        double xi[3], w, dN[60];
        or_gauss(nen, pb->reduced, q, xi, &w);
        or_shape(nen, xi, Ng[q], dN);
    }
    double M[20][20];
    for (int e = 0; e < m->n_elems; e++) {
        const double rho = pb->density[pb->params_per_elem ? e : 0];
        const int* c = m->conn + (size_t)e * nen;
        if (pb->consistent_inertia) {
            for (int a = 0; a < nen; a++)
                for (int b = 0; b < nen; b++) {
                    double s = 0.0;
                    for (int q = 0; q < ng; q++) s += m->gw[q] * rho * Ng[q][a] * Ng[q][b];
                    M[a][b] = s * m->vol[e];
                }
        } else {
            double mass = m->vol[e] * rho;
            if (mass < 4e-12) mass = 4e-12;
            mass /= nen;
            for (int a = 0; a < nen; a++)
                for (int b = 0; b < nen; b++) M[a][b] = a == b ? mass : 0.0;
        }
        for (int a = 0; a < nen; a++) {
            const int ia = m->fid[c[a]];
            if (ia < 0) continue;
            for (int b = 0; b < nen; b++) {
                if (m->fid[c[b]] < 0) continue;
                for (int d = 0; d < 3; d++) m->grav[3 * (size_t)ia + d] -= M[a][b] * g[d];
            }
        }
    }
}

or_model* or_model_create(const or_problem* pb, char* err, int errlen)
{
    int nen = pb->nen;
    if (!(nen == 4 || nen == 8 || nen == 10 || nen == 20)) {
        set_err(err, errlen, "unsupported element node count");
        return NULL;
    }
    or_model* m = (or_model*)calloc(1, sizeof(or_model));
    m->n_nodes = pb->n_nodes;
    m->n_elems = pb->n_elems;
    m->nen = nen;
    m->ng = or_n_gauss(nen, pb->reduced);
    m->law = pb->law;
    m->X = (double*)malloc(sizeof(double) * 3 * (size_t)m->n_nodes);
    memcpy(m->X, pb->X, sizeof(double) * 3 * (size_t)m->n_nodes);
    m->conn = (int*)malloc(sizeof(int) * (size_t)nen * m->n_elems);
    memcpy(m->conn, pb->conn, sizeof(int) * (size_t)nen * m->n_elems);
    for (size_t i = 0; i < (size_t)nen * m->n_elems; i++)
        if (m->conn[i] < 0 || m->conn[i] >= m->n_nodes) {
            set_err(err, errlen, "connectivity out of range");
            or_model_destroy(m);
            return NULL;
        }
    /* free-DOF numbering, Constraint.cpp:61-72 */
    m->fid = (int*)malloc(sizeof(int) * (size_t)m->n_nodes);
    int id = 0;
    for (int i = 0; i < m->n_nodes; i++)
        m->fid[i] = (pb->fixed && pb->fixed[i]) ? -1 : id++;
    m->n_free = id;
    m->ndof = 3 * id;
    m->params = (double*)malloc(sizeof(double) * 4 * (size_t)m->n_elems);
    m->coeff = (double*)malloc(sizeof(double) * (size_t)m->n_elems);
    for (int e = 0; e < m->n_elems; e++) {
        const double* p = pb->params + (pb->params_per_elem ? 4 * (size_t)e : 0);
        memcpy(m->params + 4 * (size_t)e, p, 4 * sizeof(double));
        m->coeff[e] = or_elasticity_coeff(m->law, p);
    }
    m->gw = (double*)malloc(sizeof(double) * m->ng);
    m->gdN = (double*)malloc(sizeof(double) * m->ng * nen * 3);
    for (int q = 0; q < m->ng; q++) {
        double xi[3], N[20];
        or_gauss(nen, pb->reduced, q, xi, &m->gw[q]);
        or_shape(nen, xi, N, m->gdN + (size_t)q * nen * 3);
    }
    m->vol = (double*)malloc(sizeof(double) * (size_t)m->n_elems);
    m->J0inv = (double*)malloc(sizeof(double) * 9 * (size_t)m->ng * m->n_elems);
    m->T = (double*)malloc(sizeof(double) * 3 * (size_t)nen * m->ng * m->n_elems);
    double Xe[60];
    for (int e = 0; e < m->n_elems; e++) {
        const int* c = m->conn + (size_t)e * nen;
        for (int a = 0; a < nen; a++)
            for (int d = 0; d < 3; d++)
                Xe[3 * a + d] = m->X[3 * (size_t)c[a] + d];
        m->vol[e] = pb->volumes ? pb->volumes[e] : or_element_volume(nen, Xe);
        for (int q = 0; q < m->ng; q++) {
            const double* dN = m->gdN + (size_t)q * nen * 3;
            double J0[9] = { 0 };
            for (int a = 0; a < nen; a++)
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++)
                        J0[3 * i + j] += Xe[3 * a + i] * dN[3 * a + j];
            double* Ji = m->J0inv + ((size_t)e * m->ng + q) * 9;
            inv3(J0, Ji);
            double* T = m->T + ((size_t)e * m->ng + q) * nen * 3;
            /* T = J0^-T dN^T : grad_X N_a (IntegralElement.cpp:13-17) */
            for (int a = 0; a < nen; a++)
                for (int i = 0; i < 3; i++)
                    T[3 * a + i] = Ji[i] * dN[3 * a] + Ji[3 + i] * dN[3 * a + 1] + Ji[6 + i] * dN[3 * a + 2];
        }
    }
    m->fnn = pb->fnn;
    m->pressure = 0.0;
    if (pb->n_faces > 0 && (pb->pdir || pb->pressure > 0)) {
        if (pb->pdir) {
            m->directed = 1;
            memcpy(m->pdir, pb->pdir, 3 * sizeof(double));
        } else {
            m->pressure = pb->pressure;
        }
        m->n_faces = pb->n_faces;
        m->faces = (or_face*)calloc((size_t)m->n_faces, sizeof(or_face));
        for (int f = 0; f < m->n_faces; f++)
            if (build_face(m, pb->faces + (size_t)f * pb->fnn, pb->fnn, &m->faces[f], err, errlen)) {
                or_model_destroy(m);
                return NULL;
            }
    }
    if (pb->n_fnodes > 0 && pb->fnodes && pb->fvec) {
        m->n_fnodes = pb->n_fnodes;
        m->fnodes = (int*)malloc(sizeof(int) * m->n_fnodes);
        memcpy(m->fnodes, pb->fnodes, sizeof(int) * m->n_fnodes);
        memcpy(m->fvec, pb->fvec, 3 * sizeof(double));
    }
    if (pb->gravity < 0 || pb->gravity > 6) {
        set_err(err, errlen, "gravity must be 0..6");
        or_model_destroy(m);
        return NULL;
    }
    if (pb->gravity > 0 && pb->density) build_gravity(m, pb);
    return m;
}

static void chol_free(or_chol* c)
{
    if (!c)
        return;
    free(c->perm);
    free(c->Lp);
    free(c->Li);
    free(c->Lx);
    free(c->work);
    free(c);
}

void or_model_destroy(or_model* m)
{
    if (!m)
        return;
    free(m->X);
    free(m->conn);
    free(m->fid);
    free(m->params);
    free(m->coeff);
    free(m->vol);
    free(m->gw);
    free(m->gdN);
    free(m->J0inv);
    free(m->T);
    free(m->faces);
    free(m->fnodes);
    free(m->grav);
    chol_free(m->chol);
    free(m);
}

int or_ndof(const or_model* m) { return m->ndof; }
int or_free_id(const or_model* m, int node) { return m->fid[node]; }
double or_volume_of(const or_model* m, int e) { return m->vol[e]; }
int or_face_parent(const or_model* m, int f) { return m->faces[f].elem; }

/* current position x_a = X_a + u_a (Constraint::assignDx, Constraint.cpp:93-101) */
static void node_pos(const or_model* m, const double* u, int node, double* x)
{
    int id = m->fid[node];
    for (int d = 0; d < 3; d++)
        x[d] = m->X[3 * (size_t)node + d] + (id >= 0 ? u[3 * (size_t)id + d] : 0.0);
}

/* ElementConstraint::conEvaluteResidual / conGetResidual
 * (ElementConstraint.cpp:139-162): F = J(x) J0^-1 (IntegralElement.cpp:91),
 * f_a += V w_q P(F) grad_X N_a, scattered into free DOFs. */
void or_internal_force(or_model* m, const double* u, double* r)
{
    int nen = m->nen;
    double xe[60], fe[60];
    for (int e = 0; e < m->n_elems; e++) {
        const int* c = m->conn + (size_t)e * nen;
        for (int a = 0; a < nen; a++)
            node_pos(m, u, c[a], xe + 3 * a);
        memset(fe, 0, sizeof(double) * 3 * nen);
        for (int q = 0; q < m->ng; q++) {
            const double* dN = m->gdN + (size_t)q * nen * 3;
            double J[9] = { 0 }, F[9], P[9];
            for (int a = 0; a < nen; a++)
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++)
                        J[3 * i + j] += xe[3 * a + i] * dN[3 * a + j];
            const double* Ji = m->J0inv + ((size_t)e * m->ng + q) * 9;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++)
                    F[3 * i + j] = J[3 * i] * Ji[j] + J[3 * i + 1] * Ji[3 + j] + J[3 * i + 2] * Ji[6 + j];
            or_piola(m->law, m->params + 4 * (size_t)e, F, P);
            double wv = m->vol[e] * m->gw[q];
            const double* T = m->T + ((size_t)e * m->ng + q) * nen * 3;
            for (int a = 0; a < nen; a++)
                for (int i = 0; i < 3; i++)
                    fe[3 * a + i] += wv * (P[3 * i] * T[3 * a] + P[3 * i + 1] * T[3 * a + 1] + P[3 * i + 2] * T[3 * a + 2]);
        }
        for (int a = 0; a < nen; a++) {
            int id = m->fid[c[a]];
            if (id >= 0)
                for (int d = 0; d < 3; d++)
                    r[3 * (size_t)id + d] += fe[3 * a + d];
        }
    }
}

/* follower pressure: ExternalPressureFace::getNodeDArea / getNodeArea
 * (ExternalConstraint.cpp:829-846) and ExternalPressureConstraint::
 * conEvaluteResidual / conGetResidual (:928-959); constant nodal force:
 * ExternalForceConstraint (:149-167).  Adds -f_ext into r. */
void or_external_force(or_model* m, const double* u, double* r)
{
    int nen = m->nen;
    double xe[60];
    for (int f = 0; f < m->n_faces; f++) {
        const or_face* F = &m->faces[f];
        const int* c = m->conn + (size_t)F->elem * nen;
        for (int a = 0; a < nen; a++)
            node_pos(m, u, c[a], xe + 3 * a);
        double fa[60];
        memset(fa, 0, sizeof(fa));
        for (int q = 0; q < F->nq; q++) {
            double J[9] = { 0 }, Fq[9], Fi[9], v[3];
            for (int a = 0; a < nen; a++)
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++)
                        J[3 * i + j] += xe[3 * a + i] * F->dN[q][a][j];
            const double* Ji = F->J0inv[q];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++)
                    Fq[3 * i + j] = J[3 * i] * Ji[j] + J[3 * i + 1] * Ji[3 + j] + J[3 * i + 2] * Ji[6 + j];
            double det = inv3(Fq, Fi);
            /* J F^-T A0 */
            for (int i = 0; i < 3; i++)
                v[i] = det * (Fi[i] * F->A0[0] + Fi[3 + i] * F->A0[1] + Fi[6 + i] * F->A0[2]);
            if (m->directed) {
                double nv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                for (int a = 0; a < nen; a++)
                    for (int d = 0; d < 3; d++)
                        fa[3 * a + d] += F->w[q] * F->N[q][a] * nv * m->pdir[d];
            } else {
                for (int a = 0; a < nen; a++)
                    for (int d = 0; d < 3; d++)
                        fa[3 * a + d] += F->w[q] * F->N[q][a] * v[d];
            }
        }
        for (int a = 0; a < nen; a++) {
            int id = m->fid[c[a]];
            if (id >= 0)
                for (int d = 0; d < 3; d++)
                    r[3 * (size_t)id + d] -= (m->directed ? 1.0 : m->pressure) * fa[3 * a + d];
        }
    }
    for (int i = 0; i < m->n_fnodes; i++) {
        int id = m->fid[m->fnodes[i]];
        if (id >= 0)
            for (int d = 0; d < 3; d++)
                r[3 * (size_t)id + d] -= m->fvec[d];
    }
    if (m->grav)
        for (int i = 0; i < m->ndof; i++) r[i] += m->grav[i];
}

/* Total potential (an addition): sum_e sum_q V w_q W(F_q), F = J(x) J0^-1
 * as in or_internal_force, plus u . r_c, where r_c = the constant loads'
 * part of the residual (nodal forces -f, gravity -M g), so grad = residual. */
double or_potential(or_model* m, const double* u)
{
    if (m->n_faces > 0)
        return NAN;
    int nen = m->nen;
    double xe[60], W = 0.0;
    for (int e = 0; e < m->n_elems; e++) {
        const int* c = m->conn + (size_t)e * nen;
        for (int a = 0; a < nen; a++)
            node_pos(m, u, c[a], xe + 3 * a);
        for (int q = 0; q < m->ng; q++) {
            const double* dN = m->gdN + (size_t)q * nen * 3;
            double J[9] = { 0 }, F[9];
            for (int a = 0; a < nen; a++)
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++)
                        J[3 * i + j] += xe[3 * a + i] * dN[3 * a + j];
            const double* Ji = m->J0inv + ((size_t)e * m->ng + q) * 9;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++)
                    F[3 * i + j] = J[3 * i] * Ji[j] + J[3 * i + 1] * Ji[3 + j] + J[3 * i + 2] * Ji[6 + j];
            W += m->vol[e] * m->gw[q] * or_energy_density(m->law, m->params + 4 * (size_t)e, F);
        }
    }
    double* rc = (double*)calloc((size_t)(m->ndof > 0 ? m->ndof : 1), sizeof(double));
    or_external_force(m, u, rc);
    for (int i = 0; i < m->ndof; i++)
        W += u[i] * rc[i];
    free(rc);
    return W;
}

/* Potential difference Pi(x) - Pi(xp) formed element by element,
 * sum_e sum_q V w_q (W(F_q(x)) - W(F_q(xp))) + (x - xp) . r_c, so that it
 * keeps its relative accuracy when the two states are close; *scale gets
 * sum_e sum_q V w_q |W(F_q(xp))| + |xp . r_c| (the size of Pi's rounding).
 * NAN with pressure faces. */
double or_potential_diff(or_model* m, const double* x, const double* xp, double* scale)
{
    if (m->n_faces > 0)
        return NAN;
    int nen = m->nen;
    double xa[60], xb[60], D = 0.0, S = 0.0;
    for (int e = 0; e < m->n_elems; e++) {
        const int* c = m->conn + (size_t)e * nen;
        for (int a = 0; a < nen; a++) {
            node_pos(m, x, c[a], xa + 3 * a);
            node_pos(m, xp, c[a], xb + 3 * a);
        }
        for (int q = 0; q < m->ng; q++) {
            const double* dN = m->gdN + (size_t)q * nen * 3;
            const double* Ji = m->J0inv + ((size_t)e * m->ng + q) * 9;
            double Fa[9], Fb[9];
            for (int s2 = 0; s2 < 2; s2++) {
                const double* xe = s2 ? xb : xa;
                double* F = s2 ? Fb : Fa;
                double J[9] = { 0 };
                for (int a = 0; a < nen; a++)
                    for (int i = 0; i < 3; i++)
                        for (int j = 0; j < 3; j++)
                            J[3 * i + j] += xe[3 * a + i] * dN[3 * a + j];
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++)
                        F[3 * i + j] = J[3 * i] * Ji[j] + J[3 * i + 1] * Ji[3 + j] + J[3 * i + 2] * Ji[6 + j];
            }
            const double* p = m->params + 4 * (size_t)e;
            double wa = or_energy_density(m->law, p, Fa), wb = or_energy_density(m->law, p, Fb);
            double wv = m->vol[e] * m->gw[q];
            D += wv * (wa - wb);
            S += wv * fabs(wb);
        }
    }
    double* rc = (double*)calloc((size_t)(m->ndof > 0 ? m->ndof : 1), sizeof(double));
    or_external_force(m, x, rc);
    double ext = 0.0;
    for (int i = 0; i < m->ndof; i++) {
        D += (x[i] - xp[i]) * rc[i];
        ext += xp[i] * rc[i];
    }
    free(rc);
    if (scale)
        *scale = S + fabs(ext);
    return D;
}

/* SoftObjective::residual -> Constraint::getResidual (Constraint.cpp:115-126):
 * r = f_int(u) - f_ext(u) on the free DOFs (f_ext with gravity M g when the
 * problem has it) */
void or_residual(or_model* m, const double* u, double* r)
{
    memset(r, 0, sizeof(double) * (size_t)m->ndof);
    or_internal_force(m, u, r);
    or_external_force(m, u, r);
}

/* ======================================================================== */
/* approximate Jacobian (ElementConstraint::conEvaluteJacobi,               */
/* ElementConstraint.cpp:164-205): K_e = sum_q V w_q c (T^T T) (x) I3, with  */
/* explicit zero cross-component triplets; fixed nodes dropped.             */
/* ======================================================================== */
typedef struct {
    int* p;   /* [n+1] column pointers (scalar node graph, free nodes) */
    int* i;
    double* x;
} csc_t;

static int cmp_int(const void* a, const void* b)
{
    int x = *(const int*)a, y = *(const int*)b;
    return (x > y) - (x < y);
}

/* scalar Laplacian over free nodes, full symmetric, sorted, duplicates summed */
static void scalar_laplacian(or_model* m, csc_t* L)
{
    int nf = m->n_free, nen = m->nen;
    int* cnt = (int*)calloc((size_t)nf + 1, sizeof(int));
    for (int e = 0; e < m->n_elems; e++) {
        const int* c = m->conn + (size_t)e * nen;
        for (int b = 0; b < nen; b++)
            if (m->fid[c[b]] >= 0)
                for (int a = 0; a < nen; a++)
                    if (m->fid[c[a]] >= 0)
                        cnt[m->fid[c[b]] + 1]++;
    }
    for (int j = 0; j < nf; j++)
        cnt[j + 1] += cnt[j];
    size_t tot = (size_t)cnt[nf];
    int* ri = (int*)malloc(sizeof(int) * (tot ? tot : 1));
    double* rx = (double*)malloc(sizeof(double) * (tot ? tot : 1));
    int* pos = (int*)malloc(sizeof(int) * ((size_t)nf + 1));
    memcpy(pos, cnt, sizeof(int) * ((size_t)nf + 1));
    for (int e = 0; e < m->n_elems; e++) {
        const int* c = m->conn + (size_t)e * nen;
        double Ke[400];
        memset(Ke, 0, sizeof(double) * nen * nen);
        double ks = m->vol[e] * m->coeff[e];
        for (int q = 0; q < m->ng; q++) {
            const double* T = m->T + ((size_t)e * m->ng + q) * nen * 3;
            double w = ks * m->gw[q];
            for (int a = 0; a < nen; a++)
                for (int b = 0; b < nen; b++)
                    Ke[a * nen + b] += w * (T[3 * a] * T[3 * b] + T[3 * a + 1] * T[3 * b + 1] + T[3 * a + 2] * T[3 * b + 2]);
        }
        for (int b = 0; b < nen; b++) {
            int jb = m->fid[c[b]];
            if (jb < 0)
                continue;
            for (int a = 0; a < nen; a++) {
                int ia = m->fid[c[a]];
                if (ia < 0)
                    continue;
                ri[pos[jb]] = ia;
                rx[pos[jb]] = Ke[a * nen + b];
                pos[jb]++;
            }
        }
    }
    /* sort each column and sum duplicates */
    L->p = (int*)malloc(sizeof(int) * ((size_t)nf + 1));
    L->i = (int*)malloc(sizeof(int) * (tot ? tot : 1));
    L->x = (double*)malloc(sizeof(double) * (tot ? tot : 1));
    size_t nz = 0;
    int* key = NULL;
    size_t keycap = 0;
    for (int j = 0; j < nf; j++) {
        L->p[j] = (int)nz;
        size_t b = (size_t)cnt[j], e2 = (size_t)cnt[j + 1], len = e2 - b;
        if (len > keycap) {
            keycap = 2 * len;
            key = (int*)realloc(key, sizeof(int) * 2 * keycap);
        }
        for (size_t k = 0; k < len; k++) {
            key[2 * k] = ri[b + k];
            key[2 * k + 1] = (int)(b + k);
        }
        /* sort (row, source) pairs; stable in source order within a row */
        qsort(key, len, 2 * sizeof(int), cmp_int);
        /* qsort compares the first int only; resolve ties by source index so
         * the duplicate sum runs in element order */
        for (size_t k = 0; k < len;) {
            size_t k2 = k;
            while (k2 < len && key[2 * k2] == key[2 * k])
                k2++;
            /* insertion sort the tie run by source index */
            for (size_t s = k + 1; s < k2; s++) {
                int r0 = key[2 * s], s0 = key[2 * s + 1];
                size_t t = s;
                while (t > k && key[2 * (t - 1) + 1] > s0) {
                    key[2 * t] = key[2 * (t - 1)];
                    key[2 * t + 1] = key[2 * (t - 1) + 1];
                    t--;
                }
                key[2 * t] = r0;
                key[2 * t + 1] = s0;
            }
            double s = 0;
            for (size_t t = k; t < k2; t++)
                s += rx[key[2 * t + 1]];
            L->i[nz] = key[2 * k];
            L->x[nz] = s;
            nz++;
            k = k2;
        }
    }
    L->p[nf] = (int)nz;
    free(key);
    free(cnt);
    free(ri);
    free(rx);
    free(pos);
}

long or_K_nnz(or_model* m, int scalar)
{
    csc_t L;
    scalar_laplacian(m, &L);
    long nz = L.p[m->n_free];
    free(L.p);
    free(L.i);
    free(L.x);
    return scalar ? nz : 9 * nz;
}

/* expand the scalar pattern to the reference's interleaved 3N_f matrix */
void or_K_csc(or_model* m, int scalar, int* Ap, int* Ai, double* Ax)
{
    csc_t L;
    scalar_laplacian(m, &L);
    int nf = m->n_free;
    if (scalar) {
        memcpy(Ap, L.p, sizeof(int) * ((size_t)nf + 1));
        memcpy(Ai, L.i, sizeof(int) * (size_t)L.p[nf]);
        memcpy(Ax, L.x, sizeof(double) * (size_t)L.p[nf]);
    } else {
        long nz = 0;
        for (int j = 0; j < nf; j++)
            for (int c = 0; c < 3; c++) {
                Ap[3 * j + c] = (int)nz;
                for (int k = L.p[j]; k < L.p[j + 1]; k++)
                    for (int r = 0; r < 3; r++) {
                        Ai[nz] = 3 * L.i[k] + r;
                        Ax[nz] = (r == c) ? L.x[k] : 0.0;
                        nz++;
                    }
            }
        Ap[3 * nf] = (int)nz;
    }
    free(L.p);
    free(L.i);
    free(L.x);
}

/* ======================================================================== */
/* fill-reducing ordering: geometric nested dissection over the free-node   */
/* graph (recursive median bisection of the longest axis, one-sided vertex   */
/* separator ordered last).  Any exact factorisation yields the same H0 up   */
/* to rounding; the reference's AMD only changes fill and speed.            */
/* ======================================================================== */
typedef struct {
    double key;
    int node;
} kv_t;

static int cmp_kv(const void* a, const void* b)
{
    double x = ((const kv_t*)a)->key, y = ((const kv_t*)b)->key;
    if (x < y)
        return -1;
    if (x > y)
        return 1;
    int p = ((const kv_t*)a)->node, q = ((const kv_t*)b)->node;
    return (p > q) - (p < q);
}

typedef struct {
    const int* gp;
    const int* gi;
    const double* xyz; /* [nf][3] */
    int* side;         /* scratch mark per node */
    int* out;
    int nout;
    kv_t* kv;
} nd_ctx;

static void nd_rec(nd_ctx* c, int* nodes, int n)
{
    if (n <= 32) {
        for (int k = 0; k < n; k++)
            c->out[c->nout++] = nodes[k];
        return;
    }
    double lo[3] = { 1e300, 1e300, 1e300 }, hi[3] = { -1e300, -1e300, -1e300 };
    for (int k = 0; k < n; k++)
        for (int d = 0; d < 3; d++) {
            double v = c->xyz[3 * (size_t)nodes[k] + d];
            lo[d] = v < lo[d] ? v : lo[d];
            hi[d] = v > hi[d] ? v : hi[d];
        }
    int ax = 0;
    for (int d = 1; d < 3; d++)
        if (hi[d] - lo[d] > hi[ax] - lo[ax])
            ax = d;
    for (int k = 0; k < n; k++) {
        c->kv[k].key = c->xyz[3 * (size_t)nodes[k] + ax];
        c->kv[k].node = nodes[k];
    }
    qsort(c->kv, (size_t)n, sizeof(kv_t), cmp_kv);
    int h = n / 2;
    /* side: 1 = left, 2 = right (only nodes of this set are marked) */
    for (int k = 0; k < n; k++)
        c->side[c->kv[k].node] = k < h ? 1 : 2;
    int* left = (int*)malloc(sizeof(int) * (size_t)n);
    int nl = 0, nr = 0, ns = 0;
    int* right = left + h;
    int* sep = (int*)malloc(sizeof(int) * (size_t)n);
    for (int k = 0; k < h; k++)
        left[nl++] = c->kv[k].node;
    for (int k = h; k < n; k++) {
        int v = c->kv[k].node, touch = 0;
        for (int p = c->gp[v]; p < c->gp[v + 1]; p++)
            if (c->side[c->gi[p]] == 1) {
                touch = 1;
                break;
            }
        if (touch)
            sep[ns++] = v;
        else
            right[nr++] = v;
    }
    for (int k = 0; k < n; k++)
        c->side[nodes[k]] = 0;
    if (nr == 0 || nl == 0) {
        /* no useful split: emit in sorted order */
        for (int k = 0; k < n; k++)
            c->out[c->nout++] = c->kv[k].node;
        free(left);
        free(sep);
        return;
    }
    int* tmp = (int*)malloc(sizeof(int) * (size_t)(nr > nl ? nr : nl));
    memcpy(tmp, left, sizeof(int) * (size_t)nl);
    nd_rec(c, tmp, nl);
    memcpy(tmp, right, sizeof(int) * (size_t)nr);
    nd_rec(c, tmp, nr);
    for (int k = 0; k < ns; k++)
        c->out[c->nout++] = sep[k];
    free(tmp);
    free(left);
    free(sep);
}

/* node order (new -> old) over free nodes */
static int* nd_order(or_model* m, const csc_t* L)
{
    int nf = m->n_free;
    double* xyz = (double*)malloc(sizeof(double) * 3 * ((size_t)nf + 1));
    for (int i = 0; i < m->n_nodes; i++)
        if (m->fid[i] >= 0)
            memcpy(xyz + 3 * (size_t)m->fid[i], m->X + 3 * (size_t)i, 3 * sizeof(double));
    nd_ctx c;
    c.gp = L->p;
    c.gi = L->i;
    c.xyz = xyz;
    c.side = (int*)calloc((size_t)nf + 1, sizeof(int));
    c.out = (int*)malloc(sizeof(int) * ((size_t)nf + 1));
    c.nout = 0;
    c.kv = (kv_t*)malloc(sizeof(kv_t) * ((size_t)nf + 1));
    int* nodes = (int*)malloc(sizeof(int) * ((size_t)nf + 1));
    for (int i = 0; i < nf; i++)
        nodes[i] = i;
    nd_rec(&c, nodes, nf);
    free(nodes);
    free(c.side);
    free(c.kv);
    free(xyz);
    return c.out;
}

/* ======================================================================== */
/* simplicial up-looking LL^T (the algorithm family of Eigen's SimplicialLLT */
/* used at LBFGS.h:22,70-75): elimination tree, row patterns by tree reach,  */
/* column-by-column numeric factorisation; fails on a non-positive pivot     */
/* (SimplicialCholesky_impl.h:185-188).                                      */
/* ======================================================================== */

/* C = upper triangle of P A P^T (A full symmetric CSC, pinv[old] = new) */
static void permute_upper(int n, const int* Ap, const int* Ai, const double* Ax, const int* pinv,
    int** Cp, int** Ci, double** Cx)
{
    int* cnt = (int*)calloc((size_t)n + 1, sizeof(int));
    for (int j = 0; j < n; j++)
        for (int p = Ap[j]; p < Ap[j + 1]; p++) {
            int i = Ai[p];
            int ni = pinv[i], nj = pinv[j];
            if (ni <= nj)
                cnt[nj + 1]++;
        }
    for (int j = 0; j < n; j++)
        cnt[j + 1] += cnt[j];
    int* p = (int*)malloc(sizeof(int) * ((size_t)n + 1));
    memcpy(p, cnt, sizeof(int) * ((size_t)n + 1));
    size_t nz = (size_t)cnt[n];
    int* ci = (int*)malloc(sizeof(int) * (nz ? nz : 1));
    double* cx = (double*)malloc(sizeof(double) * (nz ? nz : 1));
    for (int j = 0; j < n; j++)
        for (int q = Ap[j]; q < Ap[j + 1]; q++) {
            int ni = pinv[Ai[q]], nj = pinv[j];
            if (ni <= nj) {
                ci[p[nj]] = ni;
                cx[p[nj]] = Ax[q];
                p[nj]++;
            }
        }
    free(p);
    *Cp = cnt;
    *Ci = ci;
    *Cx = cx;
}

static int* etree_upper(int n, const int* Cp, const int* Ci)
{
    int* parent = (int*)malloc(sizeof(int) * (size_t)n);
    int* anc = (int*)malloc(sizeof(int) * (size_t)n);
    for (int k = 0; k < n; k++) {
        parent[k] = -1;
        anc[k] = -1;
        for (int p = Cp[k]; p < Cp[k + 1]; p++) {
            int i = Ci[p];
            while (i != -1 && i < k) {
                int next = anc[i];
                anc[i] = k;
                if (next == -1) {
                    parent[i] = k;
                    break;
                }
                i = next;
            }
        }
    }
    free(anc);
    return parent;
}

/* nonzero pattern of row k of L (columns < k), topologically ordered in s */
static int ereach(int k, const int* Cp, const int* Ci, const int* parent, int* s, int* flag, int n)
{
    int top = n;
    flag[k] = k;
    for (int p = Cp[k]; p < Cp[k + 1]; p++) {
        int i = Ci[p];
        if (i > k)
            continue;
        int len = 0;
        for (; flag[i] != k; i = parent[i]) {
            s[len++] = i;
            flag[i] = k;
        }
        while (len > 0)
            s[--top] = s[--len];
    }
    return top;
}

static or_chol* chol_factor(int n, const int* Ap, const int* Ai, const double* Ax, const int* perm)
{
    or_chol* ch = (or_chol*)calloc(1, sizeof(or_chol));
    ch->n = n;
    ch->perm = (int*)malloc(sizeof(int) * ((size_t)n + 1));
    memcpy(ch->perm, perm, sizeof(int) * (size_t)n);
    int* pinv = (int*)malloc(sizeof(int) * ((size_t)n + 1));
    for (int k = 0; k < n; k++)
        pinv[perm[k]] = k;
    int *Cp, *Ci;
    double* Cx;
    permute_upper(n, Ap, Ai, Ax, pinv, &Cp, &Ci, &Cx);
    int* parent = etree_upper(n, Cp, Ci);
    int* s = (int*)malloc(sizeof(int) * ((size_t)n + 1));
    int* flag = (int*)malloc(sizeof(int) * ((size_t)n + 1));
    /* column counts by row patterns */
    int* cc = (int*)calloc((size_t)n + 1, sizeof(int));
    for (int k = 0; k < n; k++)
        flag[k] = -1;
    for (int k = 0; k < n; k++) {
        int top = ereach(k, Cp, Ci, parent, s, flag, n);
        for (int t = top; t < n; t++)
            cc[s[t]]++;
        cc[k]++; /* diagonal */
    }
    ch->Lp = (int*)malloc(sizeof(int) * ((size_t)n + 1));
    long tot = 0;
    for (int k = 0; k < n; k++) {
        ch->Lp[k] = (int)tot;
        tot += cc[k];
    }
    ch->Lp[n] = (int)tot;
    ch->Li = (int*)malloc(sizeof(int) * (size_t)(tot ? tot : 1));
    ch->Lx = (double*)malloc(sizeof(double) * (size_t)(tot ? tot : 1));
    int* cp = (int*)malloc(sizeof(int) * ((size_t)n + 1));
    memcpy(cp, ch->Lp, sizeof(int) * (size_t)n);
    double* x = (double*)calloc((size_t)n + 1, sizeof(double));
    for (int k = 0; k < n; k++)
        flag[k] = -1;
    int ok = 1;
    for (int k = 0; k < n && ok; k++) {
        int top = ereach(k, Cp, Ci, parent, s, flag, n);
        x[k] = 0;
        for (int p = Cp[k]; p < Cp[k + 1]; p++)
            if (Ci[p] <= k)
                x[Ci[p]] += Cx[p];
        double d = x[k];
        x[k] = 0;
        for (; top < n; top++) {
            int i = s[top];
            double lki = x[i] / ch->Lx[ch->Lp[i]];
            x[i] = 0;
            for (int p = ch->Lp[i] + 1; p < cp[i]; p++)
                x[ch->Li[p]] -= ch->Lx[p] * lki;
            d -= lki * lki;
            int p = cp[i]++;
            ch->Li[p] = k;
            ch->Lx[p] = lki;
        }
        if (d <= 0) {
            ok = 0;
            break;
        }
        int p = cp[k]++;
        ch->Li[p] = k;
        ch->Lx[p] = sqrt(d);
    }
    free(x);
    free(cp);
    free(cc);
    free(s);
    free(flag);
    free(parent);
    free(Cp);
    free(Ci);
    free(Cx);
    free(pinv);
    if (!ok) {
        chol_free(ch);
        return NULL;
    }
    ch->work = (double*)malloc(sizeof(double) * ((size_t)n + 1));
    return ch;
}

/* x = P^T L^-T L^-1 P b (SimplicialCholesky.h:155-181) */
static void chol_solve(const or_chol* ch, const double* b, double* out)
{
    int n = ch->n;
    double* x = ch->work;
    for (int k = 0; k < n; k++)
        x[k] = b[ch->perm[k]];
    for (int j = 0; j < n; j++) {
        x[j] /= ch->Lx[ch->Lp[j]];
        for (int p = ch->Lp[j] + 1; p < ch->Lp[j + 1]; p++)
            x[ch->Li[p]] -= ch->Lx[p] * x[j];
    }
    for (int j = n - 1; j >= 0; j--) {
        for (int p = ch->Lp[j] + 1; p < ch->Lp[j + 1]; p++)
            x[j] -= ch->Lx[p] * x[ch->Li[p]];
        x[j] /= ch->Lx[ch->Lp[j]];
    }
    for (int k = 0; k < n; k++)
        out[ch->perm[k]] = x[k];
}

/* LBFGSSolver::updateApproxJacobi (LBFGS.h:65-81) */
int or_factor(or_model* m, int scalar)
{
    chol_free(m->chol);
    m->chol = NULL;
    m->chol_scalar = scalar;
    int nf = m->n_free;
    if (nf == 0)
        return 1; /* jacobi.nonZeros() == 0 -> fallback */
    csc_t L;
    scalar_laplacian(m, &L);
    int* nodes = nd_order(m, &L);
    if (scalar) {
        m->chol = chol_factor(nf, L.p, L.i, L.x, nodes);
    } else {
        int n = 3 * nf;
        long nz = 9L * L.p[nf];
        int* Ap = (int*)malloc(sizeof(int) * ((size_t)n + 1));
        int* Ai = (int*)malloc(sizeof(int) * (size_t)nz);
        double* Ax = (double*)malloc(sizeof(double) * (size_t)nz);
        or_K_csc(m, 0, Ap, Ai, Ax);
        int* perm = (int*)malloc(sizeof(int) * (size_t)n);
        for (int k = 0; k < nf; k++)
            for (int c = 0; c < 3; c++)
                perm[3 * k + c] = 3 * nodes[k] + c;
        m->chol = chol_factor(n, Ap, Ai, Ax, perm);
        free(Ap);
        free(Ai);
        free(Ax);
        free(perm);
    }
    free(nodes);
    free(L.p);
    free(L.i);
    free(L.x);
    return m->chol ? 0 : 1;
}

long or_factor_nnz(const or_model* m)
{
    return m->chol ? (long)m->chol->Lp[m->chol->n] : 0;
}

void or_apply_h0(or_model* m, const double* v, double* out)
{
    int nf = m->n_free;
    if (!m->chol) {
        memcpy(out, v, sizeof(double) * (size_t)m->ndof);
        return;
    }
    if (!m->chol_scalar) {
        chol_solve(m->chol, v, out);
        return;
    }
    /* K = L (x) I3: three scalar solves on the strided components */
    double* b = (double*)malloc(sizeof(double) * 2 * ((size_t)nf + 1));
    double* y = b + nf + 1;
    for (int c = 0; c < 3; c++) {
        for (int i = 0; i < nf; i++)
            b[i] = v[3 * (size_t)i + c];
        chol_solve(m->chol, b, y);
        for (int i = 0; i < nf; i++)
            out[3 * (size_t)i + c] = y[i];
    }
    free(b);
}

/* ======================================================================== */
/* L-BFGS                                                                   */
/* ======================================================================== */
static double dot(int n, const double* a, const double* b)
{
    double s = 0;
    for (int i = 0; i < n; i++)
        s += a[i] * b[i];
    return s;
}

/* ||a||_inf, NaN if any entry is NaN (the non-finite checks of
 * Integrator.cpp:28-34 must see it) */
static double norm_inf(int n, const double* a)
{
    double s = 0;
    for (int i = 0; i < n; i++) {
        double v = fabs(a[i]);
        s = (v != v || v > s) ? v : s;
    }
    return s;
}

/* LBFGSMat::apply_Hv (LBFGSMat.h:103-131) on explicit history */
void or_two_loop(int n, int mcap, int ncorr, int ptr, const double* S, const double* Y,
    const double* ys, double theta, or_h0_fn h0, void* ctx, const double* v, double a,
    double* res)
{
    double alpha[64];
    for (int i = 0; i < n; i++)
        res[i] = a * v[i];
    int j = ptr % mcap;
    for (int i = 0; i < ncorr; i++) {
        j = (j + mcap - 1) % mcap;
        alpha[j] = dot(n, S + (size_t)j * n, res) / ys[j];
        for (int k = 0; k < n; k++)
            res[k] -= alpha[j] * Y[(size_t)j * n + k];
    }
    if (h0) {
        double* t = (double*)malloc(sizeof(double) * (size_t)n);
        h0(ctx, res, t);
        memcpy(res, t, sizeof(double) * (size_t)n);
        free(t);
    } else {
        for (int k = 0; k < n; k++)
            res[k] /= theta;
    }
    for (int i = 0; i < ncorr; i++) {
        double beta = dot(n, Y + (size_t)j * n, res) / ys[j];
        for (int k = 0; k < n; k++)
            res[k] += (alpha[j] - beta) * S[(size_t)j * n + k];
        j = (j + 1) % mcap;
    }
}

typedef struct {
    int n;
    or_resid_fn f;
    or_h0_fn h0;
    void* ctx;
    or_energy_diff_fn e; /* potential difference for the energy line search, or NULL */
} solver_fn;

/* energy line search (an addition): Armijo constant, and the relative
 * size below which a potential difference is rounding (then only the
 * projection rule decides, as in approximate-Wolfe line searches) */
#define OR_ARMIJO_C1 1e-4
#define OR_ENERGY_RTOL 1e-12

/* LBFGSSolver::solve after the initial residual (LBFGS.h:126-171) with the
 * projection line search LineSearchFEA::LineSearch (LineSearchFEA.h:37-68).
 * r holds r(x) on entry. */
static int lbfgs_loop(const solver_fn* fn, double* x, double* r, const or_lbfgs_params* prm,
    or_lbfgs_stats* st, double* trace, int trace_cap)
{
    int n = fn->n, mc = prm->m;
    size_t nn = (size_t)n;
    double* S = (double*)calloc(nn * mc, sizeof(double));
    double* Y = (double*)calloc(nn * mc, sizeof(double));
    double* ys = (double*)calloc((size_t)mc, sizeof(double));
    double* xp = (double*)malloc(sizeof(double) * nn);
    double* rp = (double*)malloc(sizeof(double) * nn);
    double* drt = (double*)malloc(sizeof(double) * nn);
    double* tmp = (double*)malloc(sizeof(double) * nn);
    int ncorr = 0, ptr = mc;
    double theta = 1.0;
    /* initial direction d = -H0 r */
    if (fn->h0) {
        fn->h0(fn->ctx, r, drt);
        for (int i = 0; i < n; i++)
            drt[i] = -drt[i];
    } else {
        for (int i = 0; i < n; i++)
            drt[i] = -r[i];
    }
    int k = 1;
    const int energy_ls = prm->line_search == 1 && fn->e;
    double tnorm = prm->x_target ? sqrt(dot(n, prm->x_target, prm->x_target)) : 0.0;
    for (;;) {
        memcpy(xp, x, sizeof(double) * nn);
        memcpy(rp, r, sizeof(double) * nn);
        double dg_init = dot(n, r, drt);
        double cur = 1.0, used = 1.0;
        int trials = 0;
        for (int it = 0; it < prm->max_linesearch; it++) {
            for (int i = 0; i < n; i++)
                x[i] = xp[i] + cur * drt[i];
            fn->f(fn->ctx, x, r);
            st->r_evals++;
            trials++;
            used = cur;
            double dg = dot(n, r, drt);
            int proj = fabs(dg) <= prm->lns_fea * fabs(dg_init);
            if (!energy_ls) {
                /* LineSearchFEA::LineSearch (LineSearchFEA.h:37-68) */
                if (proj)
                    break;
                cur *= dg * dg_init > 0 ? 1.5 : 0.5;
                continue;
            }
            /* energy rule (addition): sufficient decrease of the potential,
             * dPi <= c1 step (r . d), or a dPi within rounding of the
             * potential, and the projection rule; without the decrease the
             * step halves, otherwise the projection rule's expansion /
             * contraction */
            double scale = 0.0;
            double dpi = fn->e(fn->ctx, x, xp, &scale);
            int armijo = dpi <= OR_ARMIJO_C1 * cur * dg_init || fabs(dpi) <= OR_ENERGY_RTOL * scale;
            if (armijo && proj)
                break;
            cur *= !armijo ? 0.5 : (dg * dg_init > 0 ? 1.5 : 0.5);
        }
        st->ls_trials += trials;
        double gnorm = norm_inf(n, r);
        if (trace && k - 1 < trace_cap) {
            trace[3 * (k - 1)] = gnorm;
            trace[3 * (k - 1) + 1] = used;
            trace[3 * (k - 1) + 2] = trials;
        }
        st->iterations = k;
        st->rinf = gnorm;
        if (prm->x_target) {
            double e2 = 0.0;
            for (int i = 0; i < n; i++)
                e2 += (x[i] - prm->x_target[i]) * (x[i] - prm->x_target[i]);
            st->target_err = sqrt(e2) / tnorm;
            if (st->target_err <= prm->target_tol) {
                st->target_reached = 1;
                break;
            }
        }
        if (gnorm <= prm->epsilon || gnorm <= prm->epsilon_rel * norm_inf(n, x))
            break;
        if (prm->max_iterations != 0 && k >= prm->max_iterations)
            break;
        /* LBFGSMat::add_correction (LBFGSMat.h:81-98), no curvature guard */
        int loc = ptr % mc;
        for (int i = 0; i < n; i++) {
            S[(size_t)loc * nn + i] = x[i] - xp[i];
            Y[(size_t)loc * nn + i] = r[i] - rp[i];
        }
        double yss = dot(n, S + (size_t)loc * nn, Y + (size_t)loc * nn);
        ys[loc] = yss;
        theta = dot(n, Y + (size_t)loc * nn, Y + (size_t)loc * nn) / yss;
        if (ncorr < mc)
            ncorr++;
        ptr = loc + 1;
        or_two_loop(n, mc, ncorr, ptr, S, Y, ys, theta, fn->h0, fn->ctx, r, -1.0, drt);
        k++;
    }
    st->xinf = norm_inf(n, x);
    free(S);
    free(Y);
    free(ys);
    free(xp);
    free(rp);
    free(drt);
    free(tmp);
    return k;
}

int or_lbfgs(int n, or_resid_fn f, or_h0_fn h0, void* ctx, double* x, const or_lbfgs_params* prm,
    or_lbfgs_stats* st, double* trace, int trace_cap)
{
    return or_lbfgs_energy(n, f, NULL, h0, ctx, x, prm, st, trace, trace_cap);
}

int or_lbfgs_energy(int n, or_resid_fn f, or_energy_diff_fn e, or_h0_fn h0, void* ctx, double* x,
    const or_lbfgs_params* prm, or_lbfgs_stats* st, double* trace, int trace_cap)
{
    memset(st, 0, sizeof(*st));
    if (prm->m <= 0 || prm->m > 64 || prm->max_linesearch <= 0)
        return -1;
    double t0 = now_s();
    double* r = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    f(ctx, x, r);
    st->r_evals = 1;
    double gnorm = norm_inf(n, r);
    st->rinf = gnorm;
    /* early exit uses ||x||_2 (LBFGS.h:118) */
    if (gnorm <= prm->epsilon || gnorm <= prm->epsilon_rel * sqrt(dot(n, x, x))) {
        st->iterations = 1;
        st->xinf = norm_inf(n, x);
        free(r);
        st->solve_seconds = now_s() - t0;
        return 1;
    }
    if (prm->line_search == 1 && !e) {
        free(r);
        return -1;
    }
    solver_fn fn = { n, f, h0, ctx, e };
    int k = lbfgs_loop(&fn, x, r, prm, st, trace, trace_cap);
    free(r);
    st->solve_seconds = now_s() - t0;
    return k;
}

static void model_resid_cb(void* ctx, const double* x, double* r)
{
    or_residual((or_model*)ctx, x, r);
}

static void model_h0_cb(void* ctx, const double* v, double* out)
{
    or_apply_h0((or_model*)ctx, v, out);
}

static double model_energy_cb(void* ctx, const double* x, const double* xp, double* scale)
{
    return or_potential_diff((or_model*)ctx, x, xp, scale);
}

/* Integrator::integral -> LBFGSSolver::solve for the model (LBFGS.h:96-174) */
int or_lbfgs_solve_model(or_model* m, double* x, const or_lbfgs_params* prm, or_lbfgs_stats* st,
    double* trace, int trace_cap)
{
    memset(st, 0, sizeof(*st));
    if (prm->m <= 0 || prm->m > 64 || prm->max_linesearch <= 0)
        return -1;
    if (prm->line_search == 1 && m->n_faces > 0)
        return -1; /* the energy rule needs a potential */
    int n = m->ndof;
    double t0 = now_s();
    double* r = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    or_residual(m, x, r);
    st->r_evals = 1;
    double gnorm = norm_inf(n, r);
    st->rinf = gnorm;
    if (gnorm <= prm->epsilon || gnorm <= prm->epsilon_rel * sqrt(dot(n, x, x))) {
        st->iterations = 1;
        st->xinf = norm_inf(n, x);
        free(r);
        st->solve_seconds = now_s() - t0;
        return 1;
    }
    int have_h0 = 0;
    if (prm->use_precond) {
        double tf = now_s();
        have_h0 = or_factor(m, prm->factor_scalar) == 0;
        st->factor_seconds = now_s() - tf;
        st->precond_failed = !have_h0;
    }
    solver_fn fn = { n, model_resid_cb, have_h0 ? model_h0_cb : NULL, m,
        prm->line_search == 1 ? model_energy_cb : NULL };
    int k = lbfgs_loop(&fn, x, r, prm, st, trace, trace_cap);
    free(r);
    st->solve_seconds = now_s() - t0;
    return k;
}
