/*
 * dgrain_compat.c -- the reference's global-state operators on top of the
 * libdgrain C ABI (see include/dgrain_compat.h).
 *
 * The globals below are the reference's data contract (2DCode/MeshAttributes.h:5-42);
 * this file only declares them.  The device copy is created from them on the
 * first call and reused while the mesh pointers stay the same.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/dgrain.h"
#include "../../include/dgrain_compat.h"

extern int NumEl;
extern int NumNodes;
extern int NumEdges;
extern double *x_coord;
extern double *y_coord;
extern double *z;
extern int *EltoVert;
extern int *EdgtoEls;
extern int *EdgtoVert;
extern int *BdryPrescribed;
extern double minEdgLength;
extern double max_lambda;
extern double *zeta;
extern double *Qx;
extern double *Qy;
extern double *Fhat1dotn;
extern double *Fhat2dotn;
extern double *Fhat3dotn;
extern double *RHSZeta;
extern double *RHSQx;
extern double *RHSQy;
extern double *bzeta;
extern double *bQn;
extern const double g;
extern double *nFriction;
extern const double H0;
extern const double VELZERO;
extern int *WD;

extern void boundary_conditions(void);

static dgr_handle *cmp_h;
static const void *cmp_key[4];
static int cmp_wetdry;
static double cmp_cfl = 0.1;

static void cmp_die(dgr_status st, const char *what)
{
    int code = 0, where = 0;
    if (cmp_h && st == DGR_ERR_NAN) dgr_last_error(cmp_h, &code, &where);
    printf("%s failed: %s (code %d at %d)\n", what, dgr_status_string(st), code, where);
    exit(EXIT_FAILURE);
}

void dgrain_compat_set_wetdry(int on) { cmp_wetdry = on; dgrain_compat_reset(); }
void dgrain_compat_set_cfl(double cfl) { cmp_cfl = cfl; dgrain_compat_reset(); }

void dgrain_compat_reset(void)
{
    if (cmp_h) dgr_destroy(cmp_h);
    cmp_h = NULL;
    memset(cmp_key, 0, sizeof(cmp_key));
}

/* the handle for the current globals, created on demand */
static dgr_handle *cmp_handle(void)
{
    const void *key[4] = {x_coord, EltoVert, EdgtoEls, (const void *)(long)NumEl};
    if (cmp_h && memcmp(key, cmp_key, sizeof(key)) == 0) return cmp_h;
    dgrain_compat_reset();
    dgr_mesh_view mv = {NumEl, NumNodes, NumEdges, x_coord, y_coord, z, nFriction,
                        EltoVert, EdgtoVert, EdgtoEls, BdryPrescribed};
    dgr_params p;
    dgr_params_default(&p);
    p.g = g;
    p.h0 = H0;
    p.velzero = VELZERO;
    p.cfl = cmp_cfl;
    p.wetdry = cmp_wetdry;
    dgr_status st = dgr_create(&mv, &p, &cmp_h);
    if (st) cmp_die(st, "dgr_create");
    memcpy(cmp_key, key, sizeof(key));
    return cmp_h;
}

static dgr_handle *cmp_upload(void)
{
    dgr_handle *h = cmp_handle();
    dgr_status st = dgr_set_state(h, zeta, Qx, Qy);
    if (!st) st = dgr_set_wd(h, WD);
    if (!st && bzeta && bQn) st = dgr_set_boundary(h, bzeta, bQn);
    if (st) cmp_die(st, "upload");
    return h;
}

static void cmp_download(dgr_handle *h)
{
    dgr_status st = dgr_get_state(h, zeta, Qx, Qy);
    if (!st) st = dgr_get_wd(h, WD);
    if (st) cmp_die(st, "download");
}

/* compute2DL.c:12-413 */
void compute2DL(double time)
{
    (void)time;
    dgr_handle *h = cmp_upload();
    dgr_status st = dgr_compute_rhs(h, RHSZeta, RHSQx, RHSQy, Fhat1dotn, Fhat2dotn, Fhat3dotn, &max_lambda);
    if (st) cmp_die(st, "compute2DL");
}

/* SlopeLimiter2D.c:10-359 */
void SlopeLimiter(void)
{
    dgr_handle *h = cmp_upload();
    dgr_status st = dgr_limit(h);
    if (st) cmp_die(st, "SlopeLimiter");
    cmp_download(h);
}

/* wetDry2D.c:16-60 */
void wetDryStatus2D(void)
{
    dgr_handle *h = cmp_upload();
    dgr_status st = dgr_wetdry_status(h);
    if (st) cmp_die(st, "wetDryStatus2D");
    cmp_download(h);
}

/* wetDry2D.c:63-260 */
void PDop2D(void)
{
    dgr_handle *h = cmp_upload();
    dgr_status st = dgr_pdop(h);
    if (st) cmp_die(st, "PDop2D");
    cmp_download(h);
}

/* time_evolution.c:10-246 */
void oneTimeStep2D(double time, double dt)
{
    dgr_handle *h = cmp_upload();
    dgr_status st = dgr_one_time_step(h, time, dt);
    if (st) cmp_die(st, "oneTimeStep2D");
    cmp_download(h);
    dgr_report r;
    /* max_lambda of the stage-2 RHS, as the global holds after the reference's step */
    st = dgr_get_report(h, &r);
    if (st) cmp_die(st, "oneTimeStep2D");
    max_lambda = r.max_lambda;
}

/* snapshot files as the reference writes them: %1.15f, three values per line */
static void cmp_write(const char *prefix, const char *zname)
{
    FILE *f[3];
    char fn[3][100];
    const double *v[3] = {zeta, Qx, Qy};
    snprintf(fn[0], 100, "%s_%s.dat", prefix, zname);
    snprintf(fn[1], 100, "%s_Qx.dat", prefix);
    snprintf(fn[2], 100, "%s_Qy.dat", prefix);
    for (int i = 0; i < 3; ++i) {
        f[i] = fopen(fn[i], "w");
        if (!f[i]) { printf("cannot write %s\n", fn[i]); exit(EXIT_FAILURE); }
    }
    for (int j = 0; j < NumEl; ++j) {
        for (int k = 0; k < 3; ++k)
            for (int i = 0; i < 3; ++i) fprintf(f[i], "%1.15f\t\t", v[i][j * 3 + k]);
        for (int i = 0; i < 3; ++i) fprintf(f[i], "\n");
    }
    for (int i = 0; i < 3; ++i) fclose(f[i]);
}

/* time_evolution.c:248-404: the step loop runs on the device (dt rule and
 * FinalTime clamp included); one step per call so that the reference's
 * per-step printout and boundary_conditions() call keep their places. */
void time_evolution(double FinalTime)
{
    boundary_conditions();                                          /* :252 */
    dgr_handle *h = cmp_upload();
    dgr_status st = dgr_set_final_time(h, FinalTime);
    if (!st) st = dgr_set_time(h, 0.0, 1e-7);                       /* :254-255 */
    if (st) cmp_die(st, "time_evolution");
    cmp_write("000", "zeta");                                       /* :260-307 */
    double time = 0, dt = 1e-7;
    int fileNumber = 1;
    while (time < FinalTime) {                                       /* :311 */
        printf("dt = %e\n", dt);                                   /* :314-315 */
        printf("time = %e\n", time + dt);
        dgr_report r;
        st = dgr_run(h, 1, &r);
        if (st) cmp_die(st, "time_evolution");
        time = r.t;
        max_lambda = r.max_lambda;
        boundary_conditions();                                      /* :319 */
        st = dgr_set_boundary(h, bzeta, bQn);
        if (st) cmp_die(st, "time_evolution");
        if (time == FinalTime) {                                     /* :324-390 */
            cmp_download(h);
            char nm[16];
            snprintf(nm, sizeof nm, "%3.3d", fileNumber);
            cmp_write(nm, "H");   /* the reference's NNN_H.dat holds zeta (:367-374) */
            ++fileNumber;
        }
        dt = r.dt;                                                  /* :399-401 */
    }
    cmp_download(h);
}
