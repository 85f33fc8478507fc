// Test driver: the host layer's output writers on data read from files.
//   writers_main <nc> <W.bin: 5 x (nc+6)^3 doubles> <parts.bin: np x (3 + 41) doubles> <np>
// Writes, in the current directory, the files of iteration 7 at t = 0.25,
// dt = 1e-3, t0 = 2 (Results/Field, Results/Particles, Results/Time,
// InitialCondition/Init000).  No GPU call is made.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "update.h"

int main(int argc, char **argv)
{
    if (argc != 5) return 2;
    nc = std::atoi(argv[1]);
    nt = nc + 6;
    nc_segx = nc_segy = nc_segz = nc;
    nt_segx = nt_segy = nt_segz = nt;
    Nc_seg = nc * nc * nc;
    Nt_seg = nt * nt * nt;
    gam = 1.4;
    t0 = 2.0;
    iter = 7;
    t = 0.25;
    dt = 1e-3;
    myrank_3d = 0;
    FOR(a, 3) sta[a] = procId[a] = 0;
    const int np = std::atoi(argv[4]);
    std::vector<double> Wb(5 * (size_t)Nt_seg), pb((size_t)np * 44);
    FILE *f = std::fopen(argv[2], "rb");
    if (!f || std::fread(Wb.data(), sizeof(double), Wb.size(), f) != Wb.size()) return 3;
    std::fclose(f);
    f = std::fopen(argv[3], "rb");
    if (!f || std::fread(pb.data(), sizeof(double), pb.size(), f) != pb.size()) return 3;
    std::fclose(f);
    ptype *W[5];
    FOR(q, 5) W[q] = Wb.data() + (size_t)q * Nt_seg;
    std::vector<ptype> prim(5 * (size_t)Nt_seg);
    ptype *P[5];
    FOR(q, 5) P[q] = prim.data() + (size_t)q * Nt_seg;
    generatePrimitives(W, P[0], P[1], P[2], P[3], P[4]);
    writeField(P[0], P[1], P[2], P[3], P[4]);
    WriteReInit(W);
    std::vector<Particle> parts(np);
    count = np;
    FOR(n, np)
    {
        parts[n].tagId = n;
        fillParticleProps(parts[n], &pb[44 * (size_t)n], &pb[44 * (size_t)n + 3]);
    }
    writeParticleData(parts.data());
    return 0;
}
