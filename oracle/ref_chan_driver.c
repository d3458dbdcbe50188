/*
 * ref_chan_driver.c -- harness around the reference's ChanNet 2D operators
 * (TEST INFRASTRUCTURE ONLY; compiled by oracle/Makefile together with the
 * reference sources under $(REF)/ChanNet into oracle/_ref/librefchan2d.so).
 *
 * ChanNet is the reference's only wet/dry-enabled 2D step that builds from
 * its own sources (ChanNet/Makefile:18, "make WD=-DWDON"); 2DCode's -DWDON
 * branch names an undeclared `junc` (2DCode/compute2DL.c:203).  This file
 * plays the part of ChanNet/main.c: it defines the program constants
 * (ChanNet/main.c:16,27,30) and fills one `struct junction`
 * (ChanNet/ChannelsAndJunctions.h:90-155) from caller arrays.
 */
#include <stdlib.h>
#include <string.h>
#include "ChannelsAndJunctions.h"

const double g = 9.810000000000;
const double H0 = 1e-3;
const double VELZERO = 1e-2;

void compute2DL(struct junction *junc, double time, double *RHSZeta, double *RHSQx, double *RHSQy);
void SlopeLimiter(struct junction *junc);
void wetDryStatus2D(struct junction *junc);
void PDop2D(struct junction *junc);

/* Arrays are borrowed from the caller and must outlive the junction. */
struct junction *refchan_create(int num_el, int num_nodes, int num_edges,
                                double *x, double *y, double *z, double *n,
                                int *elto_vert, int *edg_to_vert, int *edg_to_els, int *bdry,
                                double *zeta, double *qx, double *qy,
                                double *bzeta, double *bqn, int *wd)
{
    struct junction *j = calloc(1, sizeof(struct junction));
    j->TotalNumEdges = num_edges;
    j->NumEl = num_el;
    j->NumNodes = num_nodes;
    j->x = x; j->y = y; j->z = z; j->nFriction = n;
    j->EltoVert = elto_vert; j->EdgtoVert = edg_to_vert; j->EdgtoEls = edg_to_els;
    j->BdryPrescribed = bdry;
    j->zeta = zeta; j->Qx = qx; j->Qy = qy;
    j->bzeta = bzeta; j->bQn = bqn; j->WD = wd;
    return j;
}

void refchan_destroy(struct junction *j) { free(j); }

double refchan_compute_L(struct junction *j, double *rz, double *rqx, double *rqy)
{
    compute2DL(j, 0.0, rz, rqx, rqy);
    return j->max_lambda;
}

void refchan_limiter(struct junction *j) { SlopeLimiter(j); }

/* The channel network globals of ChanNet/main.c:19-22 and the reference's
 * coupling (ChanNet/internal_BC.c:16-217) over them.  A channel here carries
 * only what internal_BC reads: its node count, A, Q, the widths, and the
 * junction edge lists of its two ends (all borrowed from the caller). */
int NumChannels, NumJunctions;
struct channel **ChannelList;
struct junction **JunctionList;
void internal_BC(void);

struct channel *refchan_channel(int num_nodes, double *A, double *Q, double *b, int n_in, int *in_edges, int n_out,
                                int *out_edges)
{
    struct channel *c = calloc(1, sizeof(struct channel));
    c->NumNodes = num_nodes;
    c->NumEl = num_nodes - 1;
    c->A = A; c->Q = Q; c->b = b;
    c->NumInflowJunctionEdges = n_in; c->InflowJunctionEdges = in_edges;
    c->NumOutflowJunctionEdges = n_out; c->OutflowJunctionEdges = out_edges;
    return c;
}

void refchan_channel_destroy(struct channel *c) { free(c); }

void refchan_internal_bc(int nchan, struct channel **chans, int njunc, struct junction **juncs)
{
    NumChannels = nchan; ChannelList = chans;
    NumJunctions = njunc; JunctionList = juncs;
    internal_BC();
}
void refchan_wd_status(struct junction *j) { wetDryStatus2D(j); }
void refchan_pdop(struct junction *j) { PDop2D(j); }
