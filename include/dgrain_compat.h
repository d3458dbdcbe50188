/*
 * dgrain_compat.h -- the reference's own operator symbols, served by
 * libdgrain_compat.so on the GPU.
 *
 * A program built from the reference's 2DCode keeps its mesh and state in
 * globals (2DCode/MeshAttributes.h:5-42) that are defined by
 * extract_mesh_attributes.c:7-26 and initialize.c:8-27.  Link those files,
 * boundary_conditions.c and main.c as they are, and link libdgrain_compat.so
 * instead of compute2DL.c, numericalFlux2D.c, 2DInnerProducts.c,
 * SlopeLimiter2D.c, wetDry2D.c, mathfunctions.c and time_evolution.c.  Each
 * call uploads the globals it reads, runs the step on the GPU and writes back
 * the globals it changes, so the caller sees the reference's data contract.
 *
 * Replaced symbols (reference file:line):
 *   void compute2DL(double time)            2DCode/compute2DL.c:12
 *   void SlopeLimiter(void)                 2DCode/SlopeLimiter2D.c:10
 *   void wetDryStatus2D(void)               2DCode/wetDry2D.c:16
 *   void PDop2D(void)                       2DCode/wetDry2D.c:63
 *   void oneTimeStep2D(double t, double dt) 2DCode/time_evolution.c:10
 *   void time_evolution(double FinalTime)   2DCode/time_evolution.c:248
 *
 * Errors: the reference prints and calls exit(EXIT_FAILURE); so do these
 * entry points (the message names the libdgrain status).
 */
#ifndef DGRAIN_COMPAT_H
#define DGRAIN_COMPAT_H

#ifdef __cplusplus
extern "C" {
#endif

void compute2DL(double time);
void SlopeLimiter(void);
void wetDryStatus2D(void);
void PDop2D(void);
void oneTimeStep2D(double time, double dt);
void time_evolution(double FinalTime);

/* Not in the reference: the 2DCode step as shipped has its wet/dry calls
 * commented out (time_evolution.c:41,149,181,233).  Passing 1 enables them in
 * oneTimeStep2D/time_evolution (ChanNet/time_evolution.c:218-427 ordering).
 * Passing a CFL sets the dt rule's constant (time_evolution.c:399, 0.1). */
void dgrain_compat_set_wetdry(int on);
void dgrain_compat_set_cfl(double cfl);
/* Releases the device copy (the next call re-creates it from the globals). */
void dgrain_compat_reset(void);

#ifdef __cplusplus
}
#endif
#endif
