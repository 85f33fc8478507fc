// FlowField.cpp -- run parameters and initial field in the reference's binary
// formats (FlowField.cpp:9-337 of the reference), for the 1x1x1 topology.
#include "FlowField.h"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <vector>

namespace {
[[noreturn]] void die(const char *what)
{
    std::fprintf(stderr, "FlowField: %s\n", what);
    std::exit(1);
}

// periodic ghost layers of one padded plane: ghost(i) = cell(((i-3) mod n) + 3)
void fill_ghosts(ptype *w)
{
    const int n = nc_segx;  // cube
    const int m = nt_segx;
    auto src = [&](int i) { return ((i - 3) % n + n) % n + 3; };
    for (int j = 0; j < m; j++)
        for (int i = 0; i < m; i++)
            for (int k = 0; k < m; k++) {
                const int jj = src(j), ii = src(i), kk = src(k);
                if (jj != j || ii != i || kk != k) w[Is(j, i, k)] = w[Is(jj, ii, kk)];
            }
}
}  // namespace

void GetParams()
{
    // InitialConditions/params: int32 N, then gam, rho0, Mt, Re, mu, p0, eddy time
    std::ifstream in("InitialConditions/params", std::ios::in | std::ios::binary);
    if (!in) die("cannot open InitialConditions/params");
    int n = 0;
    double v[7];
    in.read(reinterpret_cast<char *>(&n), sizeof(int));
    in.read(reinterpret_cast<char *>(v), sizeof(v));
    if (!in || n <= 0) die("InitialConditions/params is truncated");
    for (int d = 0; d < 3; d++)
        if (procDim[d] <= 0) procDim[d] = 1;
    ::nc = n;
    ::nt = nc + 6;
    ::Nc = nc * nc * nc;
    ::Nt = nt * nt * nt;
    ::nc_segy = nc / procDim[0];
    ::nc_segx = nc / procDim[1];
    ::nc_segz = nc / procDim[2];
    ::Nc_seg = nc_segz * nc_segx * nc_segy;
    ::nt_segz = nc_segz + 6;
    ::nt_segx = nc_segx + 6;
    ::nt_segy = nc_segy + 6;
    ::Nt_seg = nt_segz * nt_segx * nt_segy;
    ::gam = v[0];
    ::K = int((5.0 - 3.0 * gam) / (gam - 1.0));
    ::den0 = v[1];
    ::Mt = v[2];
    ::Re = v[3];
    ::mu0 = v[4];
    ::p0 = v[5];
    ::t0 = v[6];
    ::dx = 2.0 * pi / nc;
    ::T0 = p0 / (R * den0);
    std::ifstream in2("InitialConditions/particleCount", std::ios::in | std::ios::binary);
    int np = 0;
    if (in2) in2.read(reinterpret_cast<char *>(&np), sizeof(int));
    ::numparticles = np;
    if (const char *e = std::getenv("DNSGKM_CFL")) dnsgkm_cfl = std::atof(e);
}

void Initialize_DHIT(ptype *W[5])
{
    char name[64];
    std::snprintf(name, sizeof(name), "InitialConditions/Init%i%i%i", procId[0], procId[1], procId[2]);
    std::ifstream in(name, std::ios::in | std::ios::binary);
    if (!in) die("cannot open the InitialConditions/Init file");
    std::vector<ptype> f(5 * (size_t)Nc_seg);  // u, v, w, rho, p in MATLAB order
    in.read(reinterpret_cast<char *>(f.data()), sizeof(ptype) * f.size());
    if (!in) die("the InitialConditions/Init file is truncated");
    const ptype *u0 = f.data(), *v0 = u0 + Nc_seg, *w0 = v0 + Nc_seg, *d0 = w0 + Nc_seg, *pr0 = d0 + Nc_seg;
    FOR(j, nc_segy)
    FOR(i, nc_segx)
    FOR(k, nc_segz)
    {
        const int I = Iis(j, i, k);
        const int J = k * nc_segy * nc_segx + i * nc_segy + j;  // MATLAB column-major (y, x, z)
        const ptype Vsqr = u0[J] * u0[J] + v0[J] * v0[J] + w0[J] * w0[J];
        W[0][I] = d0[J];
        W[1][I] = u0[J] * d0[J];
        W[2][I] = v0[J] * d0[J];
        W[3][I] = w0[J] * d0[J];
        W[4][I] = d0[J] * (pr0[J] / (d0[J] * (gam - 1.0)) + 0.5 * Vsqr);
    }
    PeriodicBC(W);
}

void PeriodicBC(ptype *W[5])
{
    for (int q = 0; q < 5; q++) fill_ghosts(W[q]);
}

void segment()
{
    ::sta[0] = procId[0] * nc_segy;
    ::sta[1] = procId[1] * nc_segx;
    ::sta[2] = procId[2] * nc_segz;
    ::end[0] = sta[0] + nt_segy;
    ::end[1] = sta[1] + nt_segx;
    ::end[2] = sta[2] + nt_segz;
}

void segmentFill(ptype *W[5], ptype *W0[5])
{
    For(j, nt_segy) For(i, nt_segx) For(k, nt_segz)
    {
        const int g = I(sta[0] + j, sta[1] + i, sta[2] + k);
        FOR(q, 5) W[q][Is(j, i, k)] = W0[q][g];
    }
}

void PrintParams()
{
    if (procId[0] + procId[1] + procId[2] != 0) return;
    std::cout << "\nGKM DNS DHIT (MI355X)\nInitializing Field .....\n______________________________\n\n"
              << "nc            =  " << nc << "\n"
              << "T0            =  " << T0 << "\n"
              << "Mt            =  " << Mt << "\n"
              << "Re            =  " << Re << "\n"
              << "dx            =  " << dx << "\n"
              << "eddy time     =  " << t0 << "\n"
              << "p0            =  " << p0 << "\n"
              << "den0          =  " << den0 << "\n"
              << "gam           =  " << gam << "\n"
              << "mu0           =  " << mu0 << "\n"
              << "K             =  " << K << "\n"
              << "CFL           =  " << cfl << "\n"
              << "Prandtl No.   =  " << PRN << "\n"
              << "numParticles  =  " << numparticles << "\n"
              << "nprocs    =  " << procDim[0] << " x " << procDim[1] << " x " << procDim[2] << "\n"
              << "\nField Initialized ! \n______________________________\n\nTIME              TKE\n"
              << "______________________________\n\n";
}
