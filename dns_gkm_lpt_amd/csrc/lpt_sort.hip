// lpt_sort.hip -- cell sort of the particles (hipCUB radix sort) so that the
// tricubic gathers of neighbouring threads read neighbouring coefficients.
#include <hipcub/hipcub.hpp>

#include "lpt_kernels.hpp"

namespace lpt {

// Sort key of a tracer's cell (y, x, z): blocks of kBy x kBx x kBz cells in
// row-major block order, row-major inside a block.  A block's 64 tracers (at
// one tracer per 32 cells) read an 11 x 11 x 35-cell stencil footprint of
// 100 KB, which stays in L2 while they run; ordered by the row-major cell
// index, tracers whose stencils share a y plane were a whole plane (8k
// tracers) apart and the lines were fetched again for each of the four y
// offsets.  Blocks of 2048 cells and at most 128 x 128 x 32 blocks
// (nc <= 1024) keep the key below 2^30.
constexpr int kBy = 8, kBx = 8, kBz = 32, kBlockBits = 11;  // 2^11 = kBy kBx kBz

__host__ __device__ inline unsigned block_count(int nc)
{
    return (unsigned)((nc + kBy - 1) / kBy) * ((nc + kBx - 1) / kBx) * ((nc + kBz - 1) / kBz);
}

__global__ void cell_keys_kernel(long np, const double* __restrict__ pos, int nc, double idx, unsigned* __restrict__ keys,
                                 int* __restrict__ perm)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    int c[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const double u = pos[3 * p + a] * idx;
        int v = u == u ? (int)floor(u) : 0;  // a dropped (NaN) tracer sorts into cell 0
        c[a] = v < 0 ? 0 : (v >= nc ? nc - 1 : v);
    }
    const unsigned nbx = (nc + kBx - 1) / kBx, nbz = (nc + kBz - 1) / kBz;
    const unsigned blk = ((unsigned)(c[0] / kBy) * nbx + (unsigned)(c[1] / kBx)) * nbz + (unsigned)(c[2] / kBz);
    const unsigned in = ((unsigned)(c[0] % kBy) * kBx + (unsigned)(c[1] % kBx)) * kBz + (unsigned)(c[2] % kBz);
    keys[p] = (blk << kBlockBits) | in;
    perm[p] = (int)p;
}

__global__ void gather_kernel(long np, const int* __restrict__ perm, const double* __restrict__ pos,
                              const int* __restrict__ tag, double* __restrict__ pos_out, int* __restrict__ tag_out)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const long q = perm[p];
    pos_out[3 * p + 0] = pos[3 * q + 0];
    pos_out[3 * p + 1] = pos[3 * q + 1];
    pos_out[3 * p + 2] = pos[3 * q + 2];
    tag_out[p] = tag[q];
}

__global__ void scatter_kernel(long np, const int* __restrict__ tag, const double* __restrict__ src, int width,
                               double* __restrict__ dst)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const long t = tag[p];
    for (int c = 0; c < width; c++) dst[t * width + c] = src[p * width + c];
}

static unsigned nblocks(long n) { return (unsigned)((n + 255) / 256); }

hipError_t sort_by_cell(State& s, const Geom& g, double dx, hipStream_t st)
{
    const long np = s.np;
    if (np <= 1) return hipSuccess;
    if (np > 0x7fffffffL) return hipErrorInvalidValue;
    int bits = 1;
    while ((1ul << bits) < ((unsigned long)block_count(g.nc) << kBlockBits)) bits++;
    unsigned *k_in = s.keys, *k_out = s.keys + np;
    int *v_in = s.perm, *v_out = s.perm + np;
    if (!s.sort_tmp) {
        size_t bytes = 0;
        hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, k_in, k_out, v_in, v_out, (int)np, 0, bits, st);
        if (e != hipSuccess) return e;
        e = hipMalloc(&s.sort_tmp, bytes);
        if (e != hipSuccess) return e;
        s.sort_tmp_bytes = bytes;
    }
    hipLaunchKernelGGL(cell_keys_kernel, dim3(nblocks(np)), dim3(256), 0, st, np, s.pos, g.nc, 1.0 / dx, k_in, v_in);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    size_t bytes = s.sort_tmp_bytes;
    e = hipcub::DeviceRadixSort::SortPairs(s.sort_tmp, bytes, k_in, k_out, v_in, v_out, (int)np, 0, bits, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(gather_kernel, dim3(nblocks(np)), dim3(256), 0, st, np, v_out, s.pos, s.tag, s.pos2, s.tag2);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(s.pos, s.pos2, sizeof(double) * 3 * np, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(s.tag, s.tag2, sizeof(int) * np, hipMemcpyDeviceToDevice, st);
    s.since_sort = 0;
    s.sampled = false;
    return e;
}

hipError_t scatter_to_tags(long np, const int* tag, const double* src, int width, double* dst, hipStream_t st)
{
    if (np <= 0) return hipSuccess;
    hipLaunchKernelGGL(scatter_kernel, dim3(nblocks(np)), dim3(256), 0, st, np, tag, src, width, dst);
    return hipGetLastError();
}

}  // namespace lpt
