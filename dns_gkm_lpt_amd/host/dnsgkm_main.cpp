// dnsgkm_main.cpp -- the driver sequence of the reference's Main.cpp
// (Main.cpp:10-49) on one GPU: MPI start-up with a 1x1x1 periodic Cartesian
// topology, GetParams, PrintParams, segment, the state arrays, the initial
// field, Evolve.  Run from a directory holding InitialConditions/.
// DNSGKM_NSTEPS=n stops after n steps; DNSGKM_PRECISION=mixed selects the
// mixed-precision path; DNSGKM_CFL sets the CFL number.
#include <mpi.h>

#include "FlowField.h"
#include "param.h"
#include "update.h"

int main(int argc, char *argv[])
{
    MPI_Init(&argc, &argv);
    MPI_Comm_size(MPI_COMM_WORLD, &nprocs);
    procDim[0] = procDim_y;
    procDim[1] = procDim_x;
    procDim[2] = procDim_z;
    int periodic[3] = {1, 1, 1};
    MPI_Comm comm3d;
    MPI_Cart_create(MPI_COMM_WORLD, 3, procDim, periodic, 1, &comm3d);
    MPI_Comm_rank(comm3d, &myrank_3d);
    MPI_Cart_coords(comm3d, myrank_3d, 3, procId);

    GetParams();
    PrintParams();
    segment();
    ptype *W[5];
    MAKE5s(W);
    Initialize_DHIT(W);
    Evolve(W, comm3d);

    MPI_Finalize();
    FOR(q, 5) delete[] W[q];
    return 0;
}
