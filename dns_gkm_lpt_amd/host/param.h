// param.h -- drop-in for the reference's src/param.h (globals, constants and
// layout macros the driver sequence of Main.cpp relies on).
//
// Same names and meanings as the reference (param.h:11-95) so Main.cpp compiles
// unchanged.  Differences: one process owns the whole periodic cube on one GPU,
// so the processor topology is 1x1x1, and the CFL number is a run-time global
// (default 0.10, the reference's compile-time value; env DNSGKM_CFL overrides).
#ifndef PARAMHEADERDEF
#define PARAMHEADERDEF

#include <cmath>

extern int nc, nt, Nc, Nt, nc_segx, nt_segx, nc_segy, nt_segy, nc_segz, numparticles, K, nt_segz, Nc_seg, Nt_seg,
    nprocs, procId[3], sta[3], end[3], procDim[3], myrank_3d;

#define procDim_y 1
#define procDim_x 1
#define procDim_z 1
#define pi atan(1.0) * 4.0
#define eps pow(10.0, -10.0)
#define PRN 0.70
#define cfl dnsgkm_cfl
#define R 287.0
#define nomp 5
#define SAVE_INTERVAL 1000

typedef double ptype;
extern ptype dx, T0, mu0, den0, p0, Mt, gam, Re;
extern double dnsgkm_cfl;

#define FOR(i, n) for (int i = 0; i < n; i++)
#define For(i, n) for (int i = 3; i < n - 3; i++)
#define FoR(i, n) for (int i = 2; i < n - 3; i++)
#define FOr(i, n) for (int i = 2; i < n - 2; i++)
#define FOR_(i, sta, end) for (int i = sta; i < end; i++)

// padded (y, x, z) index, z fastest
#define I(i, j, k) (i) * nt *nt + (j) * nt + (k)
#define Is(i, j, k) (i) * nt_segx *nt_segz + (j) * nt_segz + (k)
#define Ii(i, j, k) (i + 3) * nt *nt + (j + 3) * nt + (k + 3)
#define Iis(i, j, k) (i + 3) * nt_segx *nt_segz + (j + 3) * nt_segz + (k + 3)
#define Ic(i, j, k) (i) * nc *nc + (j) * nc + (k)
#define Ics(i, j, k) (i) * nc_segx *nc_segz + (j) * nc_segz + (k)
#define Icd(i, j, k) (i) * (nt_segz - 4) * (nt_segx - 4) + (j) * (nt_segz - 4) + (k)

#define MAKE(var) var = new ptype[Nt]
#define KILL(var) delete[] var
#define MAKE5(var) { MAKE(var[0]); MAKE(var[1]); MAKE(var[2]); MAKE(var[3]); MAKE(var[4]); }
#define KILL5(var) { KILL(var[0]); KILL(var[1]); KILL(var[2]); KILL(var[3]); KILL(var[4]); }
#define MAKEs(var) var = new ptype[Nt_seg];
#define MAKE5s(var) { MAKEs(var[0]); MAKEs(var[1]); MAKEs(var[2]); MAKEs(var[3]); MAKEs(var[4]); }
#define MAKE_(var) var = new ptype[Nc]
#define KILL_(var) delete[] var

extern ptype t0;
#define tmax 8 * t0

template <typename T>
int sign(T s)
{
    return s > 0 ? 1 : (s < 0 ? -1 : 0);
}

#endif
