/*
 * compat_driver.c -- test program for libdgrain_compat.so.
 *
 * Plays the part of the reference's host files (2DCode/main.c,
 * extract_mesh_attributes.c:7-26, initialize.c:8-27, boundary_conditions.c):
 * it defines the MeshAttributes.h globals, fills them from a binary input
 * written by tests/test_compat.py, calls the reference-named operators and
 * writes the globals back out.
 *
 *   compat_driver IN OUT rhs            compute2DL
 *   compat_driver IN OUT limit          SlopeLimiter
 *   compat_driver IN OUT steps N DT     N x oneTimeStep2D(t, DT)
 *   compat_driver IN OUT evolve T       time_evolution(T)
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/dgrain_compat.h"

int NumEl, NumNodes, NumEdges;
double *x_coord, *y_coord, *z, *nFriction;
int *EltoVert, *EdgtoEls, *EdgtoVert, *BdryPrescribed;
double minEdgLength = 9999999;
double max_lambda;
double *zeta, *Qx, *Qy;
double *Fhat1dotn, *Fhat2dotn, *Fhat3dotn;
double *RHSZeta, *RHSQx, *RHSQy;
double *bzeta, *bQn;
const double g = 9.81;
const double H0 = 1e-3;
const double VELZERO = 1e-2;
int *WD;

/* boundary_conditions.c:10-83 */
void boundary_conditions(void)
{
    for (int e = 0; e < NumEdges; ++e) {
        double qn = 0, zb = 0;
        if (BdryPrescribed[e] == 1) { qn = -0.042 / 0.914; zb = -1000; }
        if (BdryPrescribed[e] == 2) { zb = 0.296; qn = 1000000; }
        if (BdryPrescribed[e] == 3) { zb = -10000; qn = -0.1275 / 0.914; }
        bzeta[e] = zb;
        bQn[e] = qn;
    }
}

static void rd(FILE *f, void *p, size_t sz, size_t n)
{
    if (fread(p, sz, n, f) != n) { fprintf(stderr, "short input\n"); exit(2); }
}

int main(int argc, char **argv)
{
    if (argc < 4) { fprintf(stderr, "usage: compat_driver IN OUT op [args]\n"); return 2; }
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    int hdr[4];
    rd(f, hdr, sizeof(int), 4);
    NumEl = hdr[0]; NumNodes = hdr[1]; NumEdges = hdr[2];
    int wetdry = hdr[3];
    double cfl;
    rd(f, &cfl, sizeof(double), 1);
    size_t n3 = 3 * (size_t)NumEl;
    x_coord = malloc(NumNodes * 8); y_coord = malloc(NumNodes * 8); z = malloc(NumNodes * 8); nFriction = malloc(NumNodes * 8);
    EltoVert = malloc(n3 * 4); EdgtoVert = malloc(2 * (size_t)NumEdges * 4); EdgtoEls = malloc(2 * (size_t)NumEdges * 4);
    BdryPrescribed = malloc((size_t)NumEdges * 4);
    zeta = malloc(n3 * 8); Qx = malloc(n3 * 8); Qy = malloc(n3 * 8);
    Fhat1dotn = calloc(n3, 8); Fhat2dotn = calloc(n3, 8); Fhat3dotn = calloc(n3, 8);
    RHSZeta = calloc(n3, 8); RHSQx = calloc(n3, 8); RHSQy = calloc(n3, 8);
    bzeta = calloc(NumEdges, 8); bQn = calloc(NumEdges, 8);
    WD = malloc(NumEl * 4);
    rd(f, x_coord, 8, NumNodes); rd(f, y_coord, 8, NumNodes); rd(f, z, 8, NumNodes); rd(f, nFriction, 8, NumNodes);
    rd(f, EltoVert, 4, n3); rd(f, EdgtoVert, 4, 2 * (size_t)NumEdges); rd(f, EdgtoEls, 4, 2 * (size_t)NumEdges);
    rd(f, BdryPrescribed, 4, NumEdges);
    rd(f, &minEdgLength, 8, 1);
    rd(f, zeta, 8, n3); rd(f, Qx, 8, n3); rd(f, Qy, 8, n3);
    rd(f, WD, 4, NumEl);
    fclose(f);
    boundary_conditions();
    dgrain_compat_set_wetdry(wetdry);
    dgrain_compat_set_cfl(cfl);

    if (!strcmp(argv[3], "rhs")) {
        compute2DL(0.0);
    } else if (!strcmp(argv[3], "limit")) {
        SlopeLimiter();
    } else if (!strcmp(argv[3], "steps") && argc >= 6) {
        int n = atoi(argv[4]);
        double dt = atof(argv[5]);
        for (int i = 0; i < n; ++i) oneTimeStep2D(0.0, dt);
    } else if (!strcmp(argv[3], "evolve") && argc >= 5) {
        time_evolution(atof(argv[4]));
    } else {
        fprintf(stderr, "unknown op\n");
        return 2;
    }
    FILE *o = fopen(argv[2], "wb");
    if (!o) return 2;
    fwrite(zeta, 8, n3, o); fwrite(Qx, 8, n3, o); fwrite(Qy, 8, n3, o);
    fwrite(RHSZeta, 8, n3, o); fwrite(RHSQx, 8, n3, o); fwrite(RHSQy, 8, n3, o);
    fwrite(Fhat1dotn, 8, n3, o); fwrite(Fhat2dotn, 8, n3, o); fwrite(Fhat3dotn, 8, n3, o);
    fwrite(&max_lambda, 8, 1, o);
    fwrite(WD, 4, NumEl, o);
    fclose(o);
    dgrain_compat_reset();
    return 0;
}
