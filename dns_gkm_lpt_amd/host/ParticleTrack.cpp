// ParticleTrack.cpp -- particle bookkeeping (ParticleTrack.cpp:3-97 of the
// reference): position order pos = (y, x, z), ownership block of a position
// and the 8-double transfer record (tag, block, pos, U).
#include "ParticleTrack.h"

#include <iostream>

Particle::Particle() : infoSz(8), tagId(0), blockChanged(false), blockId(0), iter(0)
{
    FOR(q, infoSz) info[q] = 0;
}

void Particle::initializeParticle(ptype x0, ptype y0, ptype z0, int tag)
{
    tagId = tag;
    pos[1] = x0;
    pos[0] = y0;
    pos[2] = z0;
    identifyBlockId();
}

void Particle::identifyBlock()
{
    FOR(q, 3) block[q] = int(pos[q] * procDim[q] / (2 * pi));
    blockChanged = block[0] * procDim[1] * procDim[2] + block[1] * procDim[2] + block[2] != myrank_3d;
}

void Particle::identifyBlockId()
{
    identifyBlock();
    blockId = block[0] * procDim[1] * procDim[2] + block[1] * procDim[2] + block[2];
    if (blockId >= nprocs || blockId < 0)
        std::cout << " Block-Index breach found; BLOCK = " << blockId << ", " << myrank_3d << std::endl;
}

void Particle::collectInfo()
{
    info[0] = tagId;
    info[1] = blockId;
    info[2] = pos[0];
    info[3] = pos[1];
    info[4] = pos[2];
    info[5] = U[0];
    info[6] = U[1];
    info[7] = U[2];
}

void Particle::distributeInfo()
{
    tagId = (int)info[0];
    blockId = (int)info[1];
    pos[0] = info[2];
    pos[1] = info[3];
    pos[2] = info[4];
    U[0] = info[5];
    U[1] = info[6];
    U[2] = info[7];
}
