/*
 * dgswe_oracle.h -- CPU oracle for the 2D p=1 DG shallow-water RK step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * solver's step (2DCode/compute2DL.c, numericalFlux2D.c, 2DInnerProducts.c,
 * SlopeLimiter2D.c, wetDry2D.c, time_evolution.c, boundary_conditions.c,
 * extract_mesh_attributes.c).  It is used as the checker by tests/, by
 * __graft_entry__.smoke() and by bench.py's CPU-baseline leg, and by nothing
 * else: the product library never links, loads or calls it.
 *
 * Two arithmetic modes (orc_params.math):
 *   ORC_MATH_REF  squares and H^(4/3) through libm pow(), exactly as the
 *                 reference does when built with its own flags
 *                 (2DCode/Makefile:8, "-g -Wall -std=c99": gcc then emits real
 *                 pow calls, 2DInnerProducts.c:318,330,335).  Bit-identical to
 *                 the compiled reference.
 *   ORC_MATH_DEV  squares as x*x and H^(4/3) through orc_pow43(), the
 *                 deterministic pow the GPU library also implements.  This is
 *                 the mode the GPU must match bit-for-bit (up to the sign of
 *                 zero); it differs from ORC_MATH_REF only where glibc's pow
 *                 misrounds (about 0.08% of calls, 1 ulp each).
 *
 * Parity pins for each function are listed in the header comment of
 * dgswe_oracle.c.
 */
#ifndef DGSWE_ORACLE_H
#define DGSWE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_MATH_REF = 0, ORC_MATH_DEV = 1 };

typedef struct {
    int num_el, num_nodes, num_edges;
    const double *x, *y, *z, *n;      /* [num_nodes]; z = depth, positive down */
    const int *elto_vert;             /* [3*num_el] */
    const int *edg_to_vert;           /* [2*num_edges], v0 < v1 */
    const int *edg_to_els;            /* [2*num_edges], (el_in, el_ex) */
    const int *bdry;                  /* [num_edges] BdryPrescribed code */
} orc_mesh;

typedef struct {
    double g, h0, velzero;   /* 9.81, 1e-3, 1e-2 in the reference (initialize.c:22-27) */
    int wetdry;              /* 0: 2DCode as shipped (WD stays -1, no PDop); 1: WDON ordering */
    int math;                /* ORC_MATH_REF or ORC_MATH_DEV */
    double cfl;              /* reference hard-codes 0.1 (time_evolution.c:399) */
    double dt0;              /* reference 1e-7 (time_evolution.c:255) */
    double final_time;       /* dt clamp (time_evolution.c:400-401); +inf for step-count runs */
    double rain;             /* m/s, added to every RHSZeta DOF (extension) */
    int tide_code;           /* BdryPrescribed code whose bzeta follows the tide; 0 = none */
    double tide_datum, tide_amp, tide_period, tide_phase;
    int no_sentinels;        /* 1: prescribed edges take bzeta/bQn as given (ChanNet/compute2DL.c:154-161) */
} orc_params;

typedef struct orc_state orc_state;

/* --- lifetime ---------------------------------------------------------- */
orc_state *orc_create(const orc_mesh *mesh, const orc_params *p);
void orc_destroy(orc_state *s);
void orc_set_threads(orc_state *s, int nthreads);

/* --- state access (reference layout, el*3+k) --------------------------- */
void orc_set_state(orc_state *s, const double *zeta, const double *qx, const double *qy);
void orc_get_state(const orc_state *s, double *zeta, double *qx, double *qy);
void orc_set_wd(orc_state *s, const int *wd);
void orc_get_wd(const orc_state *s, int *wd);
void orc_get_modify(const orc_state *s, int stage, unsigned char *modify);
void orc_set_boundary(orc_state *s, const double *bzeta, const double *bqn);
void orc_get_boundary(const orc_state *s, double *bzeta, double *bqn);
void orc_get_fhat(const orc_state *s, double *f1, double *f2, double *f3);
void orc_get_rhs(const orc_state *s, double *rz, double *rqx, double *rqy);
double orc_max_lambda(const orc_state *s);
double orc_min_edge_length(const orc_state *s);
int orc_error(const orc_state *s, int *where);

/* --- operators (one call = one reference call) -------------------------- */
double orc_roe_flux(double F[3], double zeta_in, double zeta_ex, double qx_in, double qx_ex,
                    double qy_in, double qy_ex, double z_edge, double nx, double ny,
                    double local_g, double g);
void orc_inner_products(const orc_state *s, int el, double SI[9], double VI[21]);
int orc_compute_L(orc_state *s);
void orc_slope_limiter(orc_state *s, unsigned char *modify_out);
void orc_wetdry_status(orc_state *s);
void orc_pdop(orc_state *s);
/* elements whose state orc_pdop changed, summed over all its calls (test probe) */
long long orc_pdop_changed(const orc_state *s);
void orc_boundary_conditions(orc_state *s, double t);
/* Hydrograph extension: boundary code 1..3 takes zeta_b(t), qn_b(t),
 * piecewise linear through n points (t strictly increasing), held at the end
 * values outside them, evaluated at each step's start time.  n = 0 restores
 * the code's constants.  Not allowed on the tide code.  0 or -1 (bad input). */
int orc_set_hydrograph(orc_state *s, int code, int n, const double *t, const double *zb, const double *qn);
double orc_hydro_eval(int n, const double *t, const double *v, double time);
/* Green-Ampt infiltration extension (the reference's kinematic model,
 * KinField/Infiltration.c): cumulative depth F(t) of ponded infiltration;
 * orc_run takes the step's mean rate (F(t+dt) - F(t))/dt off the RHS of zeta
 * of every element not flagged dry.  K <= 0 turns it off.  0 or -1. */
int orc_set_infiltration(orc_state *s, double K, double psi, double theta_e, double s_e);
/* ChanNet's channel-junction coupling, junction side (ChanNet/internal_BC.c:16-217):
 * sets the coupled junction edges' bzeta/bQn from the channel end values and
 * returns the ends' ghost area and discharge.  0 or -1. */
int orc_couple_ends(orc_state *const *juncs, int nends, const int *end_off, const int *edge_junc,
                    const int *edge_local, const int *end_dir, const double *end_A, const double *end_Q,
                    double *ghost_A, double *ghost_Q);
/* DG-SWEM's coupling edge with DGSHED's 1-D channels, 2-D side
 * (src/couple2dgshed.F:22-81, src/channel_edge_hydro.F:72-75,
 * DGSHED/couple_domains.c:248-335): sets the coupled boundary edges'
 * bzeta/bQn from the channel ends and returns the 2-D edge values (swe) and
 * the ends' ghost area and discharge (ghost).  0 or -1. */
int orc_couple_dgshed(orc_state *s, int n, const int *edges, const int *conn, const double *chan,
                      const double *geom, double *swe, double *ghost);
double orc_ga_cumulative(double K, double P, double t);
double orc_log(double x);
double orc_log1p(double x);
int orc_one_time_step(orc_state *s, double t, double dt);

/* Step loop of time_evolution.c:311-402 for nsteps steps.  *t and *dt carry
 * the running time and next dt in and out. */
int orc_run(orc_state *s, int nsteps, double *t, double *dt);

/* --- math --------------------------------------------------------------- */
double orc_pow43(double x);   /* deterministic x^(4./3) (ORC_MATH_DEV) */
double orc_sin(double x);     /* deterministic sine used by the tide forcing */

/* --- topology (extract_mesh_attributes.c:143-361, restated in O(N)) ----- */
/* Returns num_edges (>0) or a negative error code.  Output arrays must hold
 * 2*(num_nodes+num_el-1) ints (edges) and num_nodes+num_el-1 ints (bdry). */
int orc_build_topology(int num_el, int num_nodes, const int *elto_vert,
                       const double *x, const double *y,
                       int n_in1, const int *in1, int n_in2, const int *in2,
                       int n_out, const int *out,
                       int *edg_to_vert, int *edg_to_els, int *bdry, double *min_edge_length);

#ifdef __cplusplus
}
#endif
#endif
