// output.cpp -- the reference's output files (update.cpp:1324-1422) and the
// full particle properties they print (getParticleProp, ParticleTrack.cpp:99-113).
//
// File names, line formats and value orders are the reference's, so existing
// post-processing reads them unchanged.  One process owns the whole cube
// (rank 0, sta = 0, procId = 0 0 0).  Directories are created when missing.
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <vector>

#include "dnsgkm_host.h"
#include "update.h"

namespace {
FILE *open_out(const char *path)
{
    std::error_code ec;
    std::filesystem::create_directories(std::filesystem::path(path).parent_path(), ec);
    FILE *f = std::fopen(path, "w");
    if (!f) {
        std::fprintf(stderr, "cannot write %s\n", path);
        std::exit(1);
    }
    return f;
}
}  // namespace

// rho, u, v, w, p over the padded segment (generateSplines, update.cpp:1202-1210)
void generatePrimitives(ptype *W[5], ptype *den, ptype *u, ptype *v, ptype *w, ptype *p)
{
    FOR(i, Nt_seg)
    {
        den[i] = W[0][i];
        u[i] = W[1][i] / den[i];
        v[i] = W[2][i] / den[i];
        w[i] = W[3][i] / den[i];
        const ptype Vsqr = u[i] * u[i] + v[i] * v[i] + w[i] * w[i];
        p[i] = (den[i] * (gam - 1)) * (W[4][i] / den[i] - 0.5 * Vsqr);
    }
}

void writeField(ptype *den, ptype *u, ptype *v, ptype *w, ptype *p)
{
    char path[64];
    std::snprintf(path, sizeof(path), "Results/Field/Field_%i_%i", myrank_3d, iter);
    FILE *f = open_out(path);
    FOR(j, nc_segy)
    FOR(i, nc_segx)
    FOR(k, nc_segz)
    {
        const int I = Iis(j, i, k);
        std::fprintf(f, "%d %d %d %.15e %.15e %.15e %.15e %.15e\n", sta[1] + i, sta[0] + j, sta[2] + k, den[I], u[I],
                     v[I], w[I], p[I]);
    }
    std::fclose(f);
}

void writeParticleData(Particle *particles)
{
    char path[64];
    std::snprintf(path, sizeof(path), "Results/Particles/P_%i_%i", myrank_3d, iter);
    FILE *f = open_out(path);
    auto v3 = [&](const ptype *a) { std::fprintf(f, "%.15e %.15e %.15e ", a[1], a[0], a[2]); };
    FOR(i, count)
    {
        const Particle &q = particles[i];
        std::fprintf(f, "%d ", q.tagId);
        v3(q.pos);
        std::fprintf(f, "%.15e ", q.DEN);
        std::fprintf(f, "%.15e %.15e %.15e ", q.U[0], q.U[1], q.U[2]);
        std::fprintf(f, "%.15e ", q.P);
        v3(q.gradD);
        v3(q.gradU);
        v3(q.gradV);
        v3(q.gradW);
        v3(q.gradP);
        v3(q.gradLapU);
        v3(q.gradLapV);
        v3(q.gradLapW);
        std::fprintf(f, "%.15e %.15e %.15e ", q.hessP[4], q.hessP[3], q.hessP[5]);
        std::fprintf(f, "%.15e %.15e %.15e ", q.hessP[1], q.hessP[0], q.hessP[2]);
        std::fprintf(f, "%.15e %.15e %.15e\n", q.hessP[7], q.hessP[6], q.hessP[8]);
    }
    std::fclose(f);
    if (myrank_3d == 0) {
        std::snprintf(path, sizeof(path), "Results/Time/T_%i", iter);
        FILE *tf = open_out(path);
        std::fprintf(tf, "%.15e %.15e %.15e", t, dt, t0);
        std::fclose(tf);
    }
}

void WriteReInit(ptype *W[5])
{
    std::vector<ptype> u0(Nc_seg), v0(Nc_seg), w0(Nc_seg), d0(Nc_seg), pr0(Nc_seg);
    FOR(j, nc_segy)
    FOR(i, nc_segx)
    FOR(k, nc_segz)
    {
        const int I = Iis(j, i, k);
        const int J = k * nc_segx * nc_segy + i * nc_segy + j;  // MATLAB order
        // c2p (update.cpp:844-853)
        const ptype deni = W[0][I];
        const ptype ui = W[1][I] / deni, vi = W[2][I] / deni, wi = W[3][I] / deni;
        const ptype Vsqr = (ui * ui + vi * vi + wi * wi);
        u0[J] = ui;
        v0[J] = vi;
        w0[J] = wi;
        d0[J] = deni;
        pr0[J] = (deni * (gam - 1)) * (W[4][I] / deni - 0.5 * Vsqr);
    }
    char path[64];
    std::snprintf(path, sizeof(path), "InitialCondition/Init%i%i%i", procId[0], procId[1], procId[2]);
    std::error_code ec;
    std::filesystem::create_directories("InitialCondition", ec);
    std::ofstream out(path, std::ios::out | std::ios::binary);
    const size_t size = sizeof(ptype) * Nc_seg;
    for (const auto *a : {&u0, &v0, &w0, &d0, &pr0}) out.write(reinterpret_cast<const char *>(a->data()), size);
    if (!out) {
        std::fprintf(stderr, "cannot write %s\n", path);
        std::exit(1);
    }
}

void fillParticleProps(Particle &q, const double *pos, const double *o)
{
    FOR(a, 3) q.pos[a] = pos[a];
    q.DEN = o[0];
    FOR(a, 3) q.U[a] = o[1 + a];
    q.P = o[4];
    FOR(a, 3)
    {
        q.gradD[a] = o[5 + a];
        q.gradU[a] = o[8 + a];
        q.gradV[a] = o[11 + a];
        q.gradW[a] = o[14 + a];
        q.gradP[a] = o[17 + a];
        q.gradLapU[a] = o[32 + a];
        q.gradLapV[a] = o[35 + a];
        q.gradLapW[a] = o[38 + a];
    }
    q.lapU = o[20];
    q.lapV = o[21];
    q.lapW = o[22];
    FOR(a, 9) q.hessP[a] = o[23 + a];
}

void getParticleProps(Particle *particles)
{
    if (count == 0) return;
    gkm_ctx *c = dnsgkm_context();
    std::vector<double> props((size_t)LPT_PROP_WIDTH * count), pos(3 * (size_t)count);
    dnsgkm_check(lpt_properties(c, props.data()), "getParticleProp");
    dnsgkm_check(lpt_download(c, pos.data(), nullptr), "lpt_download");
    FOR(n, count) fillParticleProps(particles[n], &pos[3 * (size_t)n], &props[(size_t)LPT_PROP_WIDTH * n]);
}
