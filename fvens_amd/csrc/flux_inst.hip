/* flux_inst.hip -- instantiates the flux kernel for one numerical flux (FVX_FLUX_ID is set
 * by the build, one object per flux so the instantiations compile in parallel) and exposes a
 * launcher selecting the reconstruction mode and the viscous variant. */
#include "flux_kernel.hpp"
#include "flux_wave.hpp"

#ifndef FVX_FLUX_ID
#error "FVX_FLUX_ID must be defined"
#endif

#define FVX_CAT2(a, b) a##b
#define FVX_CAT(a, b) FVX_CAT2(a, b)

namespace fvx {

void FVX_CAT(launch_flux_, FVX_FLUX_ID)(int mode, bool visc, int ncell, const double* u,
                                        const double* rgrad, const double* grad,
                                        const int* nbr, const int* face, const FaceGeom* fg,
                                        const double* centre, const double* area,
                                        const int* bcid, const GasParams* gpa,
                                        const BCTable* bta,
                                        double* res, double* dt, int getdt,
                                        const BatchArg& ba, int nact, int block, int xcd,
                                        hipStream_t s)
{
	const int cpb = block/4;
	const int nblk = (ncell + cpb - 1)/cpb;
#define FVX_LAUNCH(M, V) hipLaunchKernelGGL((k_flux<FVX_FLUX_ID, M, V>), dim3(nblk, nact), \
		dim3(block), 0, s, ncell, u, rgrad, grad, nbr, face, fg, centre, area, bcid, gpa, bta, \
		ba, res, dt, getdt, xcd)
	if(!visc) {
		if(mode == 0) FVX_LAUNCH(0, false);
		else if(mode == 1) FVX_LAUNCH(1, false);
		else FVX_LAUNCH(2, false);
	} else {
		if(mode == 0) FVX_LAUNCH(0, true);
		else if(mode == 1) FVX_LAUNCH(1, true);
		else FVX_LAUNCH(2, true);
	}
#undef FVX_LAUNCH
}

void FVX_CAT(launch_flux_wave_, FVX_FLUX_ID)(int mode, int ncell, int ngroup, const double* u,
                                             const double* rgrad, const int* tasks,
                                             const int2* groups, const unsigned* slots,
                                             const FaceGeom* fg, const double* centre,
                                             const double* area, const int* bcid,
                                             const GasParams* gpa, const BCTable* bta,
                                             double* res, double* dt, int getdt,
                                             const BatchArg& ba, int nact, int xcd,
                                             hipStream_t s)
{
	if(mode == 0)
		hipLaunchKernelGGL((k_flux_wave<FVX_FLUX_ID, 0>), dim3(ngroup, nact), dim3(64), 0, s, ncell,
		                   u, rgrad, tasks, groups, slots, fg, centre, area, bcid, gpa, bta, ba,
		                   res, dt, getdt, xcd);
	else
		hipLaunchKernelGGL((k_flux_wave<FVX_FLUX_ID, 1>), dim3(ngroup, nact), dim3(64), 0, s, ncell,
		                   u, rgrad, tasks, groups, slots, fg, centre, area, bcid, gpa, bta, ba,
		                   res, dt, getdt, xcd);
}

}
