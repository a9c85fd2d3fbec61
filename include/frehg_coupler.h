/* Coupling window of libfrehg_gw.so: the groundwater side of
 * frehg::coupling::Coupler (src/coupling/Coupler.hpp:84-175,
 * src/coupling/Coupler.cpp:69-133, src/coupling/Exchange.cpp).
 *
 * The reference's Coupler owns seep_accum, avail, gain (and bounce) and
 * aliases the surface's eta and dept into the groundwater solver
 * (Coupler.cpp:44-61).  Here those arrays are the ones attached with
 * frehg_gw_attach_coupling (device, 2D halo layout (ny+2)*(nx+2)); the calls
 * below run on the solver's stream and mutate them in place.  `bottom` is
 * the surface's bed elevation in the same layout (device).  Return codes and
 * frehg_last_error() are those of frehg_gw.h; a solver without coupling
 * attached gives FREHG_EINVAL.
 */
#ifndef FREHG_COUPLER_H
#define FREHG_COUPLER_H

#include "frehg_gw.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Coupler::beginWindow (src/coupling/Exchange.cpp:21-52): avail = dept and
 * gain = 0 on the interior, bounce (if attached) cleared.  The held-seepage
 * salt snapshot (:27-39) belongs to transport and is not built. */
int frehg_cpl_begin_window(frehg_gw_solver* s);

/* Coupler::applySeepage (src/coupling/Exchange.cpp:54-110): the accumulated
 * exchange depth goes onto eta (infiltration clamped at bottom, seepage onto
 * a dry surface held below min_depth); qss (device, may be NULL) receives
 * seep_accum / dt.  *applied and *discarded are volumes [m^3]
 * (CouplingAudit::applied / discarded, Coupler.hpp:58-59). */
int frehg_cpl_apply_seepage(frehg_gw_solver* s, const double* bottom, double dt, double* qss,
                            double* applied, double* discarded);

/* Coupler::reduceAudit (src/coupling/Exchange.cpp:112-128): the surface
 * deposit and the held accumulator as volumes (CouplingAudit::surfaceGain /
 * accumVolume).  With surface_volume non-NULL it also returns the volume of
 * a surface at rest, sum((eta - bottom) * dx * dy) over the interior (the
 * quantity SurfaceSolver::ownedVolume, src/swe/SurfaceSolver.cpp:831-844,
 * reports as sum(dept * dx * dy); counted from eta it includes films below
 * min_depth).  bottom may be NULL when surface_volume is NULL. */
int frehg_cpl_reduce_audit(frehg_gw_solver* s, const double* bottom, double* surface_gain,
                           double* accum_volume, double* surface_volume);

/* SurfaceSolver::updateDepth's cell-centre depth (src/swe/WetDry.cpp:24-43):
 * dept = eta - bottom, 0 where that is <= min_depth, over the whole plane.
 * For a caller whose surface model does not refresh dept itself before the
 * groundwater window (SurfaceSolver::refreshDepth, Coupler.cpp:77). */
int frehg_cpl_refresh_depth(frehg_gw_solver* s, const double* bottom);

#ifdef __cplusplus
}
#endif

#endif
