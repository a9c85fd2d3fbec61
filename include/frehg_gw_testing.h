/*
 * frehg_gw_testing.h — test exports of libfrehg_gw.so.
 *
 * No counterpart in the reference's public interface: these entry points
 * expose pieces of the device path to the test suite, so that the pins the
 * reference's unit tests hold can be checked on the GPU itself rather than
 * only through whole-step parity.  They are not part of the drop-in seams
 * of frehg_gw.h and callers of the solver never need them.
 */
#ifndef FREHG_GW_TESTING_H
#define FREHG_GW_TESTING_H

#include <stdint.h>

#include "frehg_gw.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Closures of frehg_test_vg (src/gw/VanGenuchten.hpp:35-128). */
#define FREHG_VG_THETA 0          /* waterContentFromHead(soil, h) */
#define FREHG_VG_HEAD 1           /* headFromWaterContent(soil, theta) */
#define FREHG_VG_CAPACITY 2       /* capacityFromHead(soil, h), the fused K/C form */
#define FREHG_VG_CONDUCTIVITY 3   /* conductivityFromHead(soil, h, ks) */
#define FREHG_VG_DKDTHETA 4       /* conductivityDerivative(soil, theta, ks) */
#define FREHG_VG_CONDUCTIVITY_C 5 /* K(h, ks) of the fused K/C form the predictor uses */

/* Evaluates one van Genuchten-Mualem closure of the device path on n points
 * with a GPU kernel, so the device closures can be compared with the hand
 * values of the reference's tests/unit/test_vangenuchten.cpp:43-126.
 * soil: n x 5 host doubles (alpha, n, theta_s, theta_r, aev) per point;
 * x: n host doubles (h, or theta for HEAD / DKDTHETA); ks: n host doubles
 * (read by CONDUCTIVITY, CONDUCTIVITY_C and DKDTHETA; may be NULL
 * otherwise); out: n host doubles. */
int frehg_test_vg(int closure, int64_t n, const double* soil, const double* x, const double* ks,
                  double* out);

/* The post-allocation sweep of one step on the solver's current state:
 * RichardsSolver::reallocateWaterContent (src/gw/Reallocate.cpp:122-545),
 * which the reference's tests call alone on an installed state
 * (tests/unit/test_gw.cpp:1037, declared at src/gw/RichardsSolver.hpp:527),
 * followed by the final clamp of finalizeWaterContent
 * (src/gw/Corrector.cpp:758-792, fused into the same device kernels) and the
 * moisture ghosts.  The saturation flags the classification reads are
 * rebuilt from the installed theta and h first.  dtg scales the face-flux
 * mirrors of the send walks.  The audit terms of the sweep replace the
 * solver's audit (frehg_gw_get_audit). */
int frehg_gw_test_post_allocate(frehg_gw_solver* s, double dtg);

#ifdef __cplusplus
}
#endif

#endif /* FREHG_GW_TESTING_H */
