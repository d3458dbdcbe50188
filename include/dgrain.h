/*
 * dgrain.h -- C ABI of libdgrain.so, the MI355X (gfx950) implementation of the
 * DG-RAIN 2D p=1 discontinuous-Galerkin shallow-water step.
 *
 * The reference (evalseth/DG-RAIN, directory 2DCode) keeps the mesh and the
 * state in globals (2DCode/MeshAttributes.h:5-42) and exposes the step as
 * void functions over them.  This ABI carries exactly the same arrays, with
 * the same layouts and index conventions, through a handle:
 *
 *   nodes    x, y, z, n            double[num_nodes]   (z = depth, positive down;
 *                                                        n = Manning coefficient)
 *   elements elto_vert             int[3*num_el]        0-based vertex ids
 *   edges    edg_to_vert           int[2*num_edges]     v0 < v1
 *            edg_to_els            int[2*num_edges]     (el_in, el_ex); el_ex == el_in
 *                                                        on a boundary edge
 *            bdry                  int[num_edges]       BdryPrescribed code
 *   state    zeta, qx, qy          double[3*num_el]     value at the midpoint of
 *                                                        local edge k at el*3+k
 *   wet/dry  wd                    int[num_el]          -1 unset, 0 dry, 1 wet
 *
 * The library copies everything to device memory at create / set time; the
 * caller keeps ownership of its arrays.  All functions return a dgr_status
 * (the reference prints and calls exit(); see 2DCode/compute2DL.c:245-249).
 * One handle binds one HIP stream on the current device; a handle must not
 * be used from two threads at once.
 *
 * Entry points and the reference interface each one replaces:
 *   dgr_build_topology   extract_mesh_attributes()  extract_mesh_attributes.c:143-361
 *   dgr_create           store_mesh()+initialize()  extract_mesh_attributes.c:39-141,
 *                                                   initialize.c:29-128 (allocation part)
 *   dgr_set_state/…      the zeta/Qx/Qy/WD globals  MeshAttributes.h:23-25,42
 *   dgr_set_boundary     boundary_conditions()      boundary_conditions.c:10-83
 *   dgr_compute_rhs      compute2DL(time)           compute2DL.c:12-413
 *   dgr_limit            SlopeLimiter()             SlopeLimiter2D.c:10-359
 *   dgr_wetdry_status    wetDryStatus2D()           wetDry2D.c:16-60
 *   dgr_pdop             PDop2D()                   wetDry2D.c:63-260
 *   dgr_one_time_step    oneTimeStep2D(time, dt)    time_evolution.c:10-246
 *   dgr_run              the step loop of time_evolution(FinalTime)
 *                                                   time_evolution.c:311-402
 */
#ifndef DGRAIN_H
#define DGRAIN_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    DGR_OK = 0,
    DGR_ERR_ARG = 1,        /* bad argument (null pointer, size, stage) */
    DGR_ERR_MESH = 2,       /* inconsistent mesh (compute2DL.c:51-55,63-67;
                               extract_mesh_attributes.c:252-257) */
    DGR_ERR_HIP = 3,        /* HIP runtime error */
    DGR_ERR_NAN = 4,        /* NaN flux or wet/dry failure (compute2DL.c:245-302,
                               wetDry2D.c:142-146,233-237); see dgr_last_error */
    DGR_ERR_NOMEM = 5,
    DGR_ERR_NODEV = 6,      /* no HIP device */
    DGR_ERR_UNSUPPORTED = 7,
    DGR_ERR_TOO_LARGE = 8   /* request exceeds a fixed workspace (dgr_gauge_run_noalloc) */
} dgr_status;

typedef struct dgr_handle dgr_handle;

typedef struct {
    int num_el, num_nodes, num_edges;
    const double *x, *y, *z, *n;
    const int *elto_vert;
    const int *edg_to_vert;
    const int *edg_to_els;
    const int *bdry;
} dgr_mesh_view;

typedef struct {
    double g;            /* 9.81     initialize.c:22 */
    double h0;           /* 1e-3     initialize.c:26 */
    double velzero;      /* 1e-2     initialize.c:27 */
    double cfl;          /* 0.1      time_evolution.c:399 */
    double dt0;          /* 1e-7     time_evolution.c:255 */
    double final_time;   /* dt clamp time_evolution.c:400-401 (HUGE_VAL: none) */
    double rain;         /* m/s added to every zeta DOF of the RHS (extension; 0 = reference) */
    int wetdry;          /* 0: 2DCode as shipped; 1: wet/dry on (ChanNet -DWDON ordering) */
    int tide_code;       /* BdryPrescribed code driven by the tide; 0 = none (extension) */
    double tide_datum, tide_amp, tide_period, tide_phase;
    int record_modify;   /* keep the limiter's per-element activation flags */
    int bc_sentinels;    /* 1 (default): 2DCode's prescribed-boundary sentinels
                            (compute2DL.c:152-155): bzeta < 0 or bQn > 1000 take the
                            interior value; 0: ChanNet's junction edges
                            (ChanNet/compute2DL.c:154-161) take bzeta/bQn as given */
} dgr_params;

typedef struct {
    double t;            /* time after the last step */
    double dt;           /* dt for the next step */
    double max_lambda;   /* max |eigenvalue| from the last stage-2 RHS (compute2DL.c:310-313) */
    int error_code;      /* 0 or the first error raised on the device */
    int error_where;     /* edge or element id of that error */
} dgr_report;

/* Reference defaults (2DCode constants, no extensions). */
void dgr_params_default(dgr_params *p);
const char *dgr_status_string(dgr_status s);

/* O(N) restatement of extract_mesh_attributes() (extract_mesh_attributes.c:143-361):
 * same edge numbering, EdgtoEls and BdryPrescribed rules.  Boundary lists are node
 * pairs (0-based) of the first inflow, second inflow and outflow sections of the
 * mesh file (extract_mesh_attributes.c:90-136).  Output arrays must hold
 * 2*(num_nodes+num_el-1) and num_nodes+num_el-1 ints.  Host only, no device use.
 * *num_edges_out receives NumEdges; DGR_ERR_MESH if Euler's count fails. */
dgr_status dgr_build_topology(int num_el, int num_nodes, const int *elto_vert,
                              const double *x, const double *y,
                              int n_in1, const int *in1, int n_in2, const int *in2,
                              int n_out, const int *out,
                              int *edg_to_vert, int *edg_to_els, int *bdry,
                              int *num_edges_out, double *min_edge_length);

/* A numbering of a mesh for memory locality: el_order[i] is the old number of
 * new element i (elements sorted along a Hilbert curve through their
 * centroids, ties in old order), node_order[j] the old number of new node j
 * (nodes in order of first use by the reordered elements; unused nodes last).
 * Arrays of num_el and num_nodes ints.  Host only.  No counterpart in the
 * reference, which steps a mesh in its file order: the kernels gather
 * neighbours' states and vertex data, and a mesh numbered without locality
 * (a shuffled fort.14) runs the step ~3x slower than the same mesh in this
 * order.  Apply it once when the mesh is read, mapping EdgtoEls and
 * EdgtoVert through the new numbers (each EdgtoVert pair re-sorted
 * ascending) but keeping the edge order and each edge's (el_in, el_ex)
 * order (dgrain.mesh.locality_order does): the run on the
 * renumbered mesh is then the original run, bit for bit.  (Edges numbered
 * anew from the new element order would flip some edges' orientation, and
 * the Roe flux evaluated the other way round differs in rounding.)
 * DGR_ERR_ARG on null arrays or negative counts, DGR_ERR_MESH on a vertex out
 * of range or a non-finite coordinate. */
dgr_status dgr_locality_order(int num_el, int num_nodes, const int *elto_vert, const double *x, const double *y,
                              int *el_order, int *node_order);

/* Handle lifetime.  Validates the mesh and builds the element-centric neighbour
 * tables.  Sets the reference's default boundary values (boundary_conditions.c)
 * and WD = -1 (initialize.c:126). */
dgr_status dgr_create(const dgr_mesh_view *mesh, const dgr_params *params, dgr_handle **out);
dgr_status dgr_destroy(dgr_handle *h);

/* Host-array state access in the reference layout (el*3+k). */
dgr_status dgr_set_state(dgr_handle *h, const double *zeta, const double *qx, const double *qy);
dgr_status dgr_get_state(dgr_handle *h, double *zeta, double *qx, double *qy);
dgr_status dgr_set_wd(dgr_handle *h, const int *wd);
dgr_status dgr_get_wd(dgr_handle *h, int *wd);
/* limiter activation flags of the last stage-`stage` limiter call (stage 0 or 1);
 * needs params.record_modify. */
dgr_status dgr_get_modify(dgr_handle *h, int stage, unsigned char *modify);
dgr_status dgr_set_boundary(dgr_handle *h, const double *bzeta, const double *bqn);
dgr_status dgr_set_time(dgr_handle *h, double t, double dt);
/* Hydrograph (extension; the reference's boundary_conditions.c:10-83 holds
 * codes 1-3 constant): the prescribed edges of BdryPrescribed code `code`
 * (1, 2 or 3; not the tide code) take zeta_b(t) and qn_b(t), piecewise linear
 * through npts points (t strictly increasing), held at the end values outside
 * them, evaluated at each step's start time by dgr_run.  The reference's
 * sentinels still apply (zeta_b < 0: interior zeta; qn_b > 1000: interior
 * normal discharge).  npts = 0 removes it.  dgr_adjoint_gauges returns
 * DGR_ERR_UNSUPPORTED while any hydrograph is set. */
dgr_status dgr_set_hydrograph(dgr_handle *h, int code, int npts, const double *t, const double *zeta_b,
                              const double *qn_b);
/* Green-Ampt infiltration (extension; the reference's kinematic model,
 * KinField/Infiltration.c:8-87 and its use in KinField/computeL.c:546-555,
 * which 2DCode does not have): conductivity K (m/s), suction head psi (m),
 * effective porosity theta_e and effective saturation s_e give
 * P = psi ((1 - s_e) theta_e), and the cumulative depth F(t) solves
 * F - K t - P ln(1 + F/P) = 0 (ponding from t = 0).  dgr_run subtracts the
 * step's mean rate (F(t + dt) - F(t)) / dt from every zeta DOF of the RHS of
 * each element not flagged dry (after params.rain).  K <= 0 turns it off;
 * otherwise psi > 0, 0 < theta_e <= 1 and 0 <= s_e < 1 or DGR_ERR_ARG.  The
 * rate is evaluated at each step's start by dgr_run; single operators
 * (dgr_compute_rhs, ...) see the rate of the handle's current time and dt,
 * as for the tide and hydrograph values (refreshed by this call, by
 * dgr_set_hydrograph and by dgr_set_time).  dgr_adjoint_gauges returns
 * DGR_ERR_UNSUPPORTED while it is on. */
dgr_status dgr_set_infiltration(dgr_handle *h, double K, double psi, double theta_e, double s_e);

/* compute2DL: RHS (and optionally the edge fluxes Fhat{1,2,3}dotn) of the
 * current state, reference layout; max_lambda may be NULL. */
dgr_status dgr_compute_rhs(dgr_handle *h, double *rhs_zeta, double *rhs_qx, double *rhs_qy,
                           double *fhat1, double *fhat2, double *fhat3, double *max_lambda);
dgr_status dgr_limit(dgr_handle *h);
dgr_status dgr_wetdry_status(dgr_handle *h);
dgr_status dgr_pdop(dgr_handle *h);

/* oneTimeStep2D(time, dt) with a host-given dt. */
dgr_status dgr_one_time_step(dgr_handle *h, double t, double dt);
/* nsteps steps of the time_evolution loop with the dt rule computed on the
 * device (dt = cfl*minEdgLength/max_lambda, clamped to final_time), starting
 * from the time/dt held by the handle (dgr_set_time; initially 0 and dt0). */
dgr_status dgr_run(dgr_handle *h, int nsteps, dgr_report *report);
/* The report of the last step (or of the last oneTimeStep2D) without stepping. */
dgr_status dgr_get_report(dgr_handle *h, dgr_report *report);
/* FinalTime of the dt clamp (time_evolution.c:400-401). */
dgr_status dgr_set_final_time(dgr_handle *h, double final_time);
dgr_status dgr_last_error(dgr_handle *h, int *code, int *where);

/* Kernel timing for measurement: between dgr_profile_begin and
 * dgr_profile_end every flux/RHS/update kernel launch (one per RK stage) is
 * bracketed by HIP events on the handle's stream; end synchronises and returns
 * their summed duration and count. */
dgr_status dgr_profile_begin(dgr_handle *h);
dgr_status dgr_profile_end(dgr_handle *h, double *total_ms, int *launches);

/* Kernel variants 0..DGR_NUM_VARIANTS-1 (all compute the same bits):
 *   0 = one fused kernel per stage (element-centric flux gather, inner
 *       products and RK update);
 *   1 = element-centric fluxes into Fhat planes, then the RHS/update kernel;
 *   2 = edge-centric fluxes (one Roe evaluation per edge) into Fhat planes,
 *       then the RHS/update kernel;
 *   3 = variant 2's kernels as one persistent cooperative launch per RK stage
 *       of a run step (grid barriers between the flux, RHS, node-bound and
 *       limiter phases; stage 2 also ends the step).  Single operators and
 *       taped steps use variant 2's separate launches; a run step in this
 *       variant is not captured into a hipGraph.
 * dgr_autotune times `steps` steps of each from the current state (restored
 * afterwards), keeps the fastest and reports ms/step per variant in
 * ms_per_step[DGR_NUM_VARIANTS] (may be NULL). */
#define DGR_NUM_VARIANTS 4
#define DGR_VARIANT_PERSIST 3
dgr_status dgr_set_variant(dgr_handle *h, int variant);
dgr_status dgr_get_variant(dgr_handle *h, int *variant);
/* Workgroup size of the RHS/update, node-bound and limiter kernels: 64, 128
 * or 256 threads (default 256; the edge kernels keep 256, their workgroups
 * carry the max-|lambda| fold partials).  Same bits at every size.
 * dgr_autotune_block times `steps` run steps at each size from the current
 * state (restored afterwards), keeps the fastest and reports it in *chosen and
 * ms/step per size (64, 128, 256) in ms_per_step[3] (may be NULL). */
dgr_status dgr_set_block(dgr_handle *h, int threads);
dgr_status dgr_get_block(dgr_handle *h, int *threads);
dgr_status dgr_autotune_block(dgr_handle *h, int steps, int *chosen, double *ms_per_step);
dgr_status dgr_autotune(dgr_handle *h, int steps, int *chosen, double *ms_per_step);

/* Run dgr_run's steps as replays of one captured hipGraph of the step
 * (on != 0) instead of enqueueing each launch; same kernels, same bits. */
dgr_status dgr_set_graph(dgr_handle *h, int on);

/* Ensembles (extension; no counterpart in the reference, which runs one
 * mesh per process).  n independent members, each with its own mesh, state,
 * boundary data, forcing, time and dt, advanced together: every kernel of a
 * step covers all members in one launch.  g, h0, velzero and wetdry must be
 * the same for every member.  dgr_run / dgr_autotune / dgr_set_graph /
 * dgr_set_final_time act on all members; the dgr_member_* calls address one
 * member (0-based) and the plain single-instance calls address member 0.
 * Each member's trajectory is bit-for-bit that of its own single-instance
 * handle.  dgr_run returns DGR_ERR_NAN if any member recorded an error; after
 * an error that member's state is undefined, the others are unaffected. */
dgr_status dgr_ensemble_create(int n, const dgr_mesh_view *meshes, const dgr_params *params,
                               dgr_handle **out);
dgr_status dgr_num_members(dgr_handle *h, int *n);
dgr_status dgr_member_set_state(dgr_handle *h, int member, const double *zeta, const double *qx,
                                const double *qy);
dgr_status dgr_member_get_state(dgr_handle *h, int member, double *zeta, double *qx, double *qy);
dgr_status dgr_member_set_wd(dgr_handle *h, int member, const int *wd);
dgr_status dgr_member_get_wd(dgr_handle *h, int member, int *wd);
dgr_status dgr_member_get_modify(dgr_handle *h, int member, int stage, unsigned char *modify);
dgr_status dgr_member_set_boundary(dgr_handle *h, int member, const double *bzeta, const double *bqn);
dgr_status dgr_member_get_boundary(dgr_handle *h, int member, double *bzeta, double *bqn);
dgr_status dgr_get_boundary(dgr_handle *h, double *bzeta, double *bqn);

/* ChanNet's channel-junction coupling, junction side (extension; the
 * reference's internal_BC, ChanNet/internal_BC.c:16-217), for junctions held
 * as the members of this handle (ChanNet's 2D junction step: params.wetdry = 1,
 * params.bc_sentinels = 0).  Channel end e couples to the junction edges
 * [end_off[e], end_off[e+1]) of edge_member/edge_local (member-local edge
 * numbers, EdgtoVert order); end_dir[e] = +1 for a channel's first node (an
 * inflow end: end_A/end_Q are its A[1], Q[1]), -1 for its last node (an
 * outflow end: A[2N-2], Q[2N-2]).  Each coupled edge takes
 * bzeta = A/(n len) - z_edge and bQn = +-Q/(n len) (n = the end's edge
 * count, len = the edge length), and the end's ghost values
 * (ChanNet's A[0], Q[0] or A[2N-1], Q[2N-1]) come back as
 * ghost_A[e] = sum of len H and ghost_Q[e] = +-sum of len qn over its edges,
 * from the members' current states.  The channels themselves stay with the
 * caller.  Bit-identical to the reference's internal_BC. */
dgr_status dgr_couple_channel_ends(dgr_handle *h, int nends, const int *end_off, const int *edge_member,
                                   const int *edge_local, const int *end_dir, const double *end_A,
                                   const double *end_Q, double *ghost_A, double *ghost_Q);

/* DG-SWEM's coupling edge with DGSHED's 1-D channels, 2-D side (extension:
 * this handle's members stand where DG-SWEM's mesh stands).  Replaces the
 * exchange of src/DG_timestep.F:55-59 -- GET_CHANNELS_VALUES
 * (obtain_channels_values, DGSHED/couple_domains.c:248-295),
 * POPULATE_SWE_ARRAY (src/couple2dgshed.F:22-81) and SHARE_SWE_VALUES
 * (receive_swe_values, DGSHED/couple_domains.c:297-335) -- for n coupled
 * boundary edges edge_member[i]/edge_local[i] (member-local edge numbers):
 *  - channel -> 2-D: chan[6i..] = (A, Q, b, m1, m2, z) of the channel end
 *    (conn[i] = 0, the channel's first node: A[1], Q[1] and element 0's
 *    b, m1, m2, z; conn[i] = 1, its last node: A[NumNodes], Q[NumNodes] and
 *    the last element's, Q negated) give zeta = getH(A, b, m1, m2) - z
 *    (DGSHED/constitutive_equations.c:6-24) and Qn; the edge takes
 *    bzeta = zeta and bQn = Qn / len, the exterior state CHANNEL_EDGE_HYDRO
 *    builds (src/channel_edge_hydro.F:72-75).  The edge needs a
 *    BdryPrescribed code and params.bc_sentinels = 0 (a channel surface may
 *    lie below the datum);
 *  - 2-D -> channel: the element's trace at the edge's two Gauss points
 *    (its CR basis as Compute2DInnerProducts evaluates it, weights 1) gives
 *    swe[2i] = 0.5 ZE_AVG and swe[2i+1] = 0.5 len Q_N_SUM (may be NULL), and
 *    with geom[4i..] = (b, m1, m2, z) of the channel end (conn 0: element
 *    0's; conn 1: the last node's NodalB, Nodalm1, Nodalm2, NodalZ) the end's
 *    ghost area ghost[2i] = h b + 0.5 (m1 + m2) h^2, h = swe[2i] + z, and
 *    discharge ghost[2i+1] = +-swe[2i+1] (the caller's A[0], Q[0] or
 *    A[NumNodes+1], Q[NumNodes+1]).
 * The trace uses the state before the new boundary data.  DGR_ERR_ARG if an
 * edge is no boundary edge (those edges are left unchanged). */
dgr_status dgr_couple_dgshed(dgr_handle *h, int n, const int *edge_member, const int *edge_local,
                             const int *conn, const double *chan, const double *geom, double *swe,
                             double *ghost);
dgr_status dgr_member_set_time(dgr_handle *h, int member, double t, double dt);
dgr_status dgr_member_set_hydrograph(dgr_handle *h, int member, int code, int npts, const double *t,
                                     const double *zeta_b, const double *qn_b);
dgr_status dgr_member_set_infiltration(dgr_handle *h, int member, double K, double psi, double theta_e,
                                       double s_e);
dgr_status dgr_member_report(dgr_handle *h, int member, dgr_report *report);
/* element range of a member inside the device planes of dgr_device_state */
dgr_status dgr_member_layout(dgr_handle *h, int member, int *el_offset, int *num_el);

/* Gradients (extension; no counterpart in the reference).  A gauge-misfit
 * loss over a forward run from the handle's current state,
 *     L = sum over samples s, gauges g of (h_g(s) - obs[s][g])^2,
 * with h_g the mean depth (zeta + z at the midpoints) of gauge element g after
 * every `sample_every`-th step, and its gradient with respect to the nodal
 * depth z and Manning n of each member, by a hand-written reverse of the RK
 * step (Roe flux, inner products, Manning friction, RK update, limiter with
 * its active set, dt = cfl*minEdgLength/max|lambda| and the tide's time) over
 * checkpoints every `checkpoint_every` steps.  The forward's wet/dry flags
 * enter as recorded, and PDop2D's mass and momentum corrections are
 * differentiated on the branch each element took.  The final-time clamp of dt
 * is taken as inactive.
 * The handle's state, flags and time are restored on return. */
typedef struct {
    int nsteps;
    int sample_every;
    int checkpoint_every;
    int ngauges;              /* gauges per member */
    const int *gauge_el;      /* [members][ngauges] member-local element ids */
    const double *obs;        /* [members][nsteps/sample_every][ngauges] */
} dgr_adjoint_spec;
/* loss[members]; grad_z, grad_n: [sum of members' num_nodes], member order (may be NULL) */
dgr_status dgr_adjoint_gauges(dgr_handle *h, const dgr_adjoint_spec *spec, double *loss, double *grad_z,
                              double *grad_n);
/* Device memory dgr_adjoint_gauges may spend on resident tapes (extension).
 * The last R segments (as many as fit) are taped by the forward pass itself
 * and reversed without recomputation; the others are recomputed from their
 * checkpoints.  bytes < 0 (the default): the device memory still free when
 * the call has allocated its other buffers, less 1/16 of the device's
 * memory; 0: recompute every segment.  The gradient is the same bit for bit.
 * The tape space stays allocated on the handle for the next call (a gradient
 * loop repeats the same sizes) until this call or dgr_destroy releases it. */
dgr_status dgr_set_tape_budget(dgr_handle *h, long long bytes);
/* Forward run that samples the gauge depths: h_out[members][nsteps/sample_every][ngauges]. */
dgr_status dgr_gauge_run(dgr_handle *h, int nsteps, int sample_every, int ngauges, const int *gauge_el,
                         double *h_out);
/* dgr_gauge_run with no device allocation (extension; no counterpart in the
 * reference).  For callers that cannot allocate device memory mid-run:
 * serving loops and harnesses that forbid hipMalloc.  The gauge ids and the
 * samples are staged in a fixed per-device workspace of
 * DGR_NOALLOC_SCRATCH_BYTES that each device holds until the process ends:
 * the ids rounded up to 256 bytes plus members*ngauges*(nsteps/sample_every)
 * doubles must fit, else DGR_ERR_TOO_LARGE.  Same results as dgr_gauge_run.
 * All calls on one device share the workspace: calls that overlap on one
 * device (two threads, two handles) race. */
#define DGR_NOALLOC_SCRATCH_BYTES (8u << 20)
dgr_status dgr_gauge_run_noalloc(dgr_handle *h, int nsteps, int sample_every, int ngauges, const int *gauge_el,
                                 double *h_out);
/* Replace a member's nodal depth z and/or Manning n (NULL keeps it).  The
 * handle's derived bed data (edge geometry, per-node depth plane) follow, so
 * the next steps are those of a handle created on the new bed. */
dgr_status dgr_set_nodal(dgr_handle *h, int member, const double *z, const double *n);

/* Restart files (extension): one member's state, wet/dry flags, time and dt.
 * Layout: "DGRCKPT1", int32 num_el, double t, double dt, zeta/Qx/Qy
 * (3*num_el doubles each, el*3+k), WD (num_el int32), host byte order.
 * Reading checks the tag and num_el (DGR_ERR_ARG otherwise). */
dgr_status dgr_checkpoint_write(dgr_handle *h, int member, const char *path);
dgr_status dgr_checkpoint_read(dgr_handle *h, int member, const char *path);

/* Synchronise the handle's stream. */
dgr_status dgr_sync(dgr_handle *h);
/* The handle's HIP stream (a hipStream_t), on which every kernel of the
 * handle runs: callers time it with their own HIP events, or order their own
 * work after the handle's. */
dgr_status dgr_get_stream(dgr_handle *h, void **stream);

/* Device-resident state: pointer to the 9 SoA planes [zeta0..2, qx0..2, qy0..2]
 * of `pitch` doubles each (pitch >= num_el).  Valid until dgr_destroy. */
dgr_status dgr_device_state(dgr_handle *h, double **planes, int *pitch);

/* Self-test of the device math against host-supplied inputs: out[i] =
 * {sqrt(x), x/y, dgr_pow43(x), dgr_sin(x), fma(x,y,x), x/y and x/3 by the
 * shared-denominator division, ln(1 + |x|), the Green-Ampt depth F(t = |x|)
 * for K = 1e-6 + 1e-5 |y|, P = 0.033} for each pair (9 doubles). */
dgr_status dgr_math_selftest(int n, const double *x, const double *y, double *out);

#ifdef __cplusplus
}
#endif
#endif
