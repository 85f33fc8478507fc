// update.h -- drop-in for the reference's src/update.h.
//
// Evolve(W, comm3d) (update.h:37) runs the reference's time loop
// (update.cpp:39-188) with the state resident on one GPU through the C ABI of
// include/dnsgkm.h.  The other entry points keep the reference's meaning on
// host arrays in the reference's padded layout; the splines they took in the
// reference live on the device inside the solver context.
#ifndef UPDATEHEADERDEF
#define UPDATEHEADERDEF

#include <mpi.h>

#include "ParticleTrack.h"
#include "param.h"

extern ptype dt, t;
extern int iter, count;

void Evolve(ptype *W[5], MPI_Comm comm3d);
// one Eulerian step of size dt on the host state W (GPU evolve, GPUupdate.cu:79-219)
void evolveEulerian(ptype *W[5]);
// dt from the CFL condition of W (update.cpp:1031-1054)
void timestep(ptype *W[5]);
// periodic ghost layers of W (HaloTransfer, update.cpp:855-1029, single domain)
void HaloTransfer(ptype *W[5]);
// the statistics line of PrintStats (update.cpp:688-842) for W and the particles
void PrintStats(ptype *W[5]);
// particle positions from InitialConditions/unirand (locateParticles, update.cpp:1286-1322)
void locateParticles(Particle *particles);
// sample the particles' velocity and gradients on the current field (getParticleProp)
void sampleParticles(Particle *particles);
// particle statistics of the last sample (GetParticleStats, update.cpp:1111-1162)
void GetParticleStats(Particle *particles);
// every member of every particle on the current field (getParticleProp,
// ParticleTrack.cpp:99-113, with the eight splines of generateSplines)
void getParticleProps(Particle *particles);
// one particle's members from pos (y, x, z) and an lpt_properties row
void fillParticleProps(Particle &q, const double *pos, const double *props);

// output files (output.cpp; update.cpp:1324-1422)
void generatePrimitives(ptype *W[5], ptype *den, ptype *u, ptype *v, ptype *w, ptype *p);
void writeField(ptype *den, ptype *u, ptype *v, ptype *w, ptype *p);
void writeParticleData(Particle *particles);
void WriteReInit(ptype *W[5]);

#endif
