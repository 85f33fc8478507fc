// ParticleTrack.h -- drop-in for the reference's src/ParticleTrack.h.
//
// The Particle record keeps the reference's members (ParticleTrack.h:14-36)
// and its bookkeeping methods.  Sampling and Heun advection of particles run
// for all particles at once on the GPU inside Evolve (lpt_coupled_step), so
// the per-particle spline methods of the reference (track, updatePosition,
// getParticleProp, which take einspline objects) have batched counterparts
// in update.h instead.
#ifndef PARTICLETRACKHEADERDEF
#define PARTICLETRACKHEADERDEF

#include "param.h"

class Particle
{
public:
    Particle();
    int infoSz, tagId;
    ptype info[8];
    bool blockChanged;
    ptype pos[3], DEN, U[3], P, gradD[3], gradU[3], gradV[3], gradW[3], gradP[3], lapU, lapV, lapW, hessP[9],
        gradLapU[3], gradLapV[3], gradLapW[3];

    int blockId, iter, block[3];
    void initializeParticle(ptype x0, ptype y0, ptype z0, int tag);
    void identifyBlock();
    void identifyBlockId();
    void collectInfo();
    void distributeInfo();
};

#endif
