// Host check of the compensated update's carry layout (gkm_kernels.hpp,
// carry_plane / carry_index): for each grid size given, every cell (y, x, z)
// of nc^3 maps to its own index below carry_plane(nc), and the 16 cells of a
// z block of one line are 16 consecutive floats.
//   carry_layout_check nc...     exit 0 when every size passes
// Built with -DPLANT_DROP_BLOCK the z block number is left out of the index
// (cells 16 apart along z collide), which the check must reject.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "gkm_kernels.hpp"

static long index_of(int y, int x, int z, int nc)
{
#ifdef PLANT_DROP_BLOCK
    return (((long)x * nc + y) * 16) + (z & 15);
#else
    return gkm::carry_index(y, x, z, nc);
#endif
}

int main(int argc, char** argv)
{
    int bad = 0;
    for (int i = 1; i < argc; i++) {
        const int nc = atoi(argv[i]);
        const long cp = gkm::carry_plane(nc);
        std::vector<unsigned char> seen((size_t)cp, 0);
        long out_of_range = 0, repeated = 0, not_contiguous = 0;
        for (int y = 0; y < nc; y++)
            for (int x = 0; x < nc; x++)
                for (int z = 0; z < nc; z++) {
                    const long k = index_of(y, x, z, nc);
                    if (k < 0 || k >= cp) { out_of_range++; continue; }
                    if (seen[(size_t)k]++) repeated++;
                    if ((z & 15) && k != index_of(y, x, z - 1, nc) + 1) not_contiguous++;
                }
        printf("nc %d  carry_plane %ld (nc^3 %ld)  out of range %ld  repeated %ld  not contiguous %ld\n", nc, cp,
               (long)nc * nc * nc, out_of_range, repeated, not_contiguous);
        if (out_of_range || repeated || not_contiguous) bad = 1;
    }
    return bad;
}
