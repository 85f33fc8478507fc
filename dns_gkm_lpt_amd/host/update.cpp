// update.cpp -- the reference's time loop (update.cpp:39-188) driving the
// resident GPU solver through the C ABI.
//
// Per iteration, as the reference orders it: every SAVE_INTERVAL iterations the
// field and restart files (writeField, WriteReInit); every 10th iteration the
// particle properties from the field at t^n (getParticleProp), the particle
// file (writeParticleData), the particle statistics and the statistics line
// (PrintStats); the Eulerian step; the periodic halo; the velocity spline of
// t^{n+1}; the Heun particle step; the new CFL dt.  On the GPU all of it after
// the first upload stays on the device; the host sees the statistics, the
// output files and, at the end, the state and the particles.  DNSGKM_OUTPUT=0
// turns the output files off.
#include "update.h"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "dnsgkm_host.h"

ptype dt, t;
int iter, count;

namespace {
gkm_ctx *g_ctx = nullptr;
MPI_Comm g_comm = MPI_COMM_WORLD;
gkm_pstats g_pstats{};
bool g_have_pstats = false;
}  // namespace

gkm_ctx *dnsgkm_context()
{
    if (g_ctx) return g_ctx;
    gkm_config cfg{};
    cfg.nc = nc;
    cfg.precision = GKM_FP64;
    if (const char *e = std::getenv("DNSGKM_PRECISION"))
        if (std::string(e) == "mixed") cfg.precision = GKM_MIXED;
    cfg.gam = gam;
    cfg.rho0 = den0;
    cfg.Mt = Mt;
    cfg.Re = Re;
    cfg.mu0 = mu0;
    cfg.p0 = p0;
    cfg.t0 = t0;
    cfg.cfl = cfl;
    cfg.prandtl = PRN;
    cfg.device = 0;
    cfg.check_state = 1;
    gkm_ctx *c = nullptr;
    int rc = gkm_create(&cfg, &c);
    if (rc) {
        std::fprintf(stderr, "gkm_create: %s\n", c ? gkm_last_error(c) : "failed");
        std::exit(1);
    }
    g_ctx = c;
    return g_ctx;
}

void dnsgkm_check(int rc, const char *what)
{
    if (rc == GKM_OK) return;
    std::fprintf(stderr, "%s: %s\n", what, gkm_last_error(g_ctx));
    std::exit(1);
}

void evolveEulerian(ptype *W[5])
{
    gkm_ctx *c = dnsgkm_context();
    dnsgkm_check(gkm_upload_state(c, W), "evolveEulerian upload");
    dnsgkm_check(gkm_step(c, dt), "evolveEulerian step");
    dnsgkm_check(gkm_download_state(c, W), "evolveEulerian download");
}

void timestep(ptype *W[5])
{
    gkm_ctx *c = dnsgkm_context();
    dnsgkm_check(gkm_upload_state(c, W), "timestep upload");
    dnsgkm_check(gkm_cfl_dt(c, &dt), "timestep");
}

void HaloTransfer(ptype *W[5])
{
    gkm_ctx *c = dnsgkm_context();
    dnsgkm_check(gkm_upload_state(c, W), "HaloTransfer upload");
    dnsgkm_check(gkm_download_state(c, W), "HaloTransfer download");
}

void locateParticles(Particle *particles)
{
    count = 0;
    if (numparticles <= 0) return;
    std::ifstream in("InitialConditions/unirand", std::ios::in | std::ios::binary);
    std::vector<ptype> xyz(3 * (size_t)numparticles);
    if (in) in.read(reinterpret_cast<char *>(xyz.data()), sizeof(ptype) * xyz.size());
    if (!in) {
        std::fprintf(stderr, "locateParticles: cannot read %d positions from InitialConditions/unirand\n",
                     numparticles);
        std::exit(1);
    }
    std::vector<double> pos(3 * (size_t)numparticles);
    FOR(q, numparticles)
    {
        particles[q].initializeParticle(xyz[q], xyz[numparticles + q], xyz[2 * numparticles + q], q);
        FOR(a, 3) pos[3 * (size_t)q + a] = particles[q].pos[a];
    }
    count = numparticles;
    dnsgkm_check(lpt_upload(dnsgkm_context(), numparticles, pos.data()), "lpt_upload");
}

void sampleParticles(Particle *particles)
{
    if (count == 0) return;
    std::vector<double> U(3 * (size_t)count), gU(U.size()), gV(U.size()), gW(U.size());
    dnsgkm_check(lpt_sample(dnsgkm_context(), U.data(), gU.data(), gV.data(), gW.data()), "lpt_sample");
    FOR(q, count)
    FOR(a, 3)
    {
        particles[q].U[a] = U[3 * (size_t)q + a];
        particles[q].gradU[a] = gU[3 * (size_t)q + a];
        particles[q].gradV[a] = gV[3 * (size_t)q + a];
        particles[q].gradW[a] = gW[3 * (size_t)q + a];
    }
}

void GetParticleStats(Particle *)
{
    g_have_pstats = count > 0 && lpt_stats(dnsgkm_context(), &g_pstats) == GKM_OK;
}

void PrintStats(ptype *[5])
{
    gkm_stats s{};
    dnsgkm_check(gkm_stats_compute(dnsgkm_context(), &s), "PrintStats");
    int rank = 0;
    MPI_Comm_rank(g_comm, &rank);
    if (rank != 0) return;
    const gkm_pstats p = g_have_pstats ? g_pstats : gkm_pstats{0, 0, 0, 0};
    std::cout.precision(8);
    std::cout << std::scientific << t / t0 << " " << s.tke << " " << s.diss << " " << s.sijsij << " " << s.skew[0]
              << " " << s.skew[1] << " " << s.skew[2] << " " << s.kurt[0] << " " << s.kurt[1] << " " << s.kurt[2]
              << " " << s.prms << " " << s.trms << " "
              << " " << p.tke << " "
              << " " << p.sijsij << " " << p.skew_u << " " << p.skew_w << std::endl;
}

void Evolve(ptype *W[5], MPI_Comm comm3d)
{
    g_comm = comm3d;
    gkm_ctx *c = dnsgkm_context();
    dnsgkm_check(gkm_upload_state(c, W), "Evolve upload");
    Particle *particles = new Particle[numparticles > 0 ? numparticles : 1];
    locateParticles(particles);
    dnsgkm_check(gkm_cfl_dt(c, &dt), "timestep");
    t = 0;
    iter = 0;
    long nmax = -1;  // a fixed step count instead of t <= tmax
    if (const char *e = std::getenv("DNSGKM_NSTEPS")) nmax = std::atol(e);
    bool output = true;
    if (const char *e = std::getenv("DNSGKM_OUTPUT")) output = std::string(e) != "0";
    while (t <= tmax && (nmax < 0 || iter < nmax)) {
        if (output && iter % SAVE_INTERVAL == 0) {
            dnsgkm_check(gkm_download_state(c, W), "writeField download");
            std::vector<ptype> prim(5 * (size_t)Nt_seg);
            ptype *P[5];
            FOR(q, 5) P[q] = prim.data() + (size_t)q * Nt_seg;
            generatePrimitives(W, P[0], P[1], P[2], P[3], P[4]);
            writeField(P[0], P[1], P[2], P[3], P[4]);
            WriteReInit(W);
        }
        if (iter % 10 == 0) {
            if (count > 0) {
                if (output) {
                    getParticleProps(particles);
                    writeParticleData(particles);
                }
                dnsgkm_check(lpt_sample(c, nullptr, nullptr, nullptr, nullptr), "getParticleProp");
                GetParticleStats(particles);
            }
            PrintStats(W);
        }
        if (count > 0)
            dnsgkm_check(lpt_coupled_step(c, dt), "Evolve step");
        else
            dnsgkm_check(gkm_step(c, dt), "Evolve step");
        dnsgkm_check(gkm_cfl_dt(c, &dt), "timestep");
        t += dt;
        iter++;
    }
    dnsgkm_check(gkm_download_state(c, W), "Evolve download");
    if (count > 0) {
        std::vector<double> pos(3 * (size_t)count);
        dnsgkm_check(lpt_download(c, pos.data(), nullptr), "lpt_download");
        FOR(q, count) FOR(a, 3) particles[q].pos[a] = pos[3 * (size_t)q + a];
    }
    delete[] particles;
}
