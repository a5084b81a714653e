/*
 * ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * Drives the reference's own CPU sources, compiled where they lie by oracle/Makefile
 * (src/global.cpp, src/opt/fobj.cpp, src/opt/linesearch.cpp, src/flux_kinetic.cpp,
 * src/flux_moment.cpp, src/comm.cpp serial branch), into oracle/_ref/libref.so.
 * It only sets the reference's globals (src/global.h:65-129) and forwards calls; it
 * replaces nothing the reference needs.  The rest of the reference (opt.cpp,
 * optBatch.cpp, utils.cpp, init.cpp) needs LAPACK and GSL, which this image lacks, so
 * it is not built.
 *
 * Several reference functions keep first-call static buffers sized by the first grid
 * (flux_kinetic.cpp:70-80, flux_moment.cpp:32-51), so tests load a fresh copy of the
 * library per grid configuration.
 */
#include <stddef.h>
#include "global.h"
#include "opt/opt.h"

extern "C" {

void ref_setup(int nx, int ny, int nz, int nm, int nm2, int nq, double theta, double *w,
               double *xi, double *eta, double *mu, double *M, double *M2, double *beta,
               double *betaVal, int *betaColInd, int *betaRowPtr, double *pnOps)
{
    g_nx = nx;
    g_ny = ny;
    g_nz = nz;
    g_gX[0] = 0; g_gX[1] = NUM_GHOST_CELLS; g_gX[2] = NUM_GHOST_CELLS + nx - 1; g_gX[3] = 2 * NUM_GHOST_CELLS + nx - 1;
    g_gY[0] = 0; g_gY[1] = NUM_GHOST_CELLS; g_gY[2] = NUM_GHOST_CELLS + ny - 1; g_gY[3] = 2 * NUM_GHOST_CELLS + ny - 1;
    g_gZ[0] = 0; g_gZ[1] = NUM_GHOST_CELLS; g_gZ[2] = NUM_GHOST_CELLS + nz - 1; g_gZ[3] = 2 * NUM_GHOST_CELLS + nz - 1;
    g_numMoments = nm;
    g_numMoments2 = nm2;
    g_numQuadPoints = nq;
    g_theta = theta;
    g_w = w;
    g_xi = xi;
    g_eta = eta;
    g_mu = mu;
    g_m = M;
    g_m2 = M2;
    g_beta = beta;
    g_betaVal = betaVal;
    g_betaColInd = betaColInd;
    g_betaRowPtr = betaRowPtr;
    if (pnOps) {
        size_t s = (size_t)nm * nm;
        g_pnFluxOperatorXiPlus = pnOps + 0 * s;
        g_pnFluxOperatorXiMinus = pnOps + 1 * s;
        g_pnFluxOperatorEtaPlus = pnOps + 2 * s;
        g_pnFluxOperatorEtaMinus = pnOps + 3 * s;
        g_pnFluxOperatorMuPlus = pnOps + 4 * s;
        g_pnFluxOperatorMuMinus = pnOps + 5 * s;
    }
    g_cudaKinetic = false;
    g_node = 0;
    fobj_init(nm2);
    linesearch_init(nm);
}

void ref_solveHMn(double *alpha, double *h, int useGaunt, int useGauntSparse)
{
    solveHMn(g_numMoments, g_numMoments2, g_numQuadPoints, alpha, g_w, g_m, g_m2, h,
             useGaunt != 0, useGauntSparse != 0);
}

double ref_fobjMn(double *alpha, double *u, double *g, double *h, int useGaunt,
                  int useGauntSparse)
{
    return fobjMn(g_numMoments, g_numMoments2, g_numQuadPoints, alpha, u, g_w, g_m, g_m2, g, h,
                  useGaunt != 0, useGauntSparse != 0);
}

double ref_fobjMnUseH(double *alpha, double *u, double *g, double *h)
{
    return fobjMnUseH(g_numMoments, alpha, u, g_m, g, h);
}

double ref_linesearch(double *alpha, double *d, double f, double *g, double *u)
{
    return linesearch(g_numMoments, g_numMoments2, g_numQuadPoints, alpha, d, f, g, u, g_w, g_m);
}

void ref_solveFluxKinetic(double *u, double *flux, double *alpha, double dx, double dy,
                          double dz, int isPn)
{
    g_solver = isPn ? SOLVER_PN : SOLVER_MN;
    solveFluxKinetic(u, flux, alpha, dx, dy, dz);
}

void ref_solveFluxMoment(double *u, double *flux, double dx, double dy, double dz)
{
    solveFluxMoment(u, flux, dx, dy, dz);
}

void ref_communicateBoundaries(double *u)
{
    g_u = u;
    communicateBoundaries();
}

} /* extern "C" */
