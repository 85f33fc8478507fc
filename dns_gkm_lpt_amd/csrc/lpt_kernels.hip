// lpt_kernels.hip -- natural tricubic B-spline prefilter, particle sampling and
// Heun advection, and flow / particle statistics on gfx950.
//
// Spline semantics (einspline UBspline_3d_d with NATURAL ends as built by
// get_bsplinep, Interpolation.cpp:3-31): nt = nc+6 knots per axis at the
// padded cell centres, N = nt+2 coefficients per axis, interpolation rows
// (c_i + 4 c_{i+1} + c_{i+2})/6 = f_i and f'' = 0 at both ends, i.e.
// c_1 = f_0, c_nt = f_{nt-1}, a [1 4 1] system for the inner coefficients and
// c_0 = 2c_1 - c_2, c_{nt+1} = 2c_nt - c_{nt-1}.  The 3-D coefficients are the
// tensor product: the line solve is applied along z, then x, then y.
#include "lpt_kernels.hpp"
#include "lpt_device.hpp"

#include <initializer_list>
#include <type_traits>

namespace lpt {

void release(State& s)
{
    (void)hipFree(s.pos);
    (void)hipFree(s.U);
    (void)hipFree(s.G);
    for (int b = 0; b < 2; b++)
        for (int q = 0; q < 3; q++) {
            (void)hipFree(s.coef[b][q]);
            s.coef[b][q] = nullptr;
        }
    for (int f = 0; f < 5; f++) (void)hipFree(s.pcoef[f]);
    (void)hipFree((void*)s.pcs);
    (void)hipFree(s.props);
    (void)hipFree(s.pivots);
    (void)hipFree(s.red);
    (void)hipFree(s.tag);
    (void)hipFree(s.tag2);
    (void)hipFree(s.pos2);
    (void)hipFree(s.keys);
    (void)hipFree(s.perm);
    (void)hipFree(s.sort_tmp);
    s = State();
}

// four waves per block, one z line each
template <int K>
__global__ __launch_bounds__(256) void prefilter_z_kernel(ZArgs z)
{
    const long line = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (line >= (long)z.g.nt * z.g.nt) return;  // uniform over the wave
    prefilter_z_line<K>(line, z);
}

// z pass in place over knot values already in the coefficient box (positions
// [0, nt) of every z line with a, b < nt), as the Laplacian splines need
template <int K>
__global__ __launch_bounds__(256) void knots_z_kernel(double* __restrict__ coef, Geom g, const double* __restrict__ piv)
{
    const long line = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int nt = g.nt, lane = threadIdx.x & 63;
    if (line >= (long)nt * nt) return;
    double* row = coef + ((line / nt) * g.N + line % nt) * g.N;
    double f[K];
#pragma unroll
    for (int j = 0; j < K; j++) {
        const int p = 64 * j + lane;
        f[j] = p < nt ? row[p] : 0.0;
    }
    zline_solve<K>(f, nt, piv, row);  // every lane has read its values before the first store
}

// smallest instantiated chunk count holding nt knots (0: too many)
static int z_chunks(int nt)
{
    const int need = (nt + 63) / 64;
    for (int k : {1, 2, 3, 5, 9, kMaxZChunks})
        if (need <= k) return k;
    return 0;
}

template <typename F>
static hipError_t with_chunks(int k, F&& f)
{
    switch (k) {
    case 1: return f(std::integral_constant<int, 1>());
    case 2: return f(std::integral_constant<int, 2>());
    case 3: return f(std::integral_constant<int, 3>());
    case 5: return f(std::integral_constant<int, 5>());
    case 9: return f(std::integral_constant<int, 9>());
    case kMaxZChunks: return f(std::integral_constant<int, kMaxZChunks>());
    default: return hipErrorInvalidValue;
    }
}

// up to three coefficient boxes of one geometry, solved by one launch per axis
struct CoefSet {
    double* c[3];
};

// Workgroups [f bpf, (f + 1) bpf) solve the lines of box f (bpf = workgroups
// per box).  A workgroup is 256 consecutive lines, so its four waves (one per
// SIMD of a CU) read and write each row as one 2 KB run rather than four 512 B
// runs of unrelated workgroups: the passes are bound by how HBM serves runs at
// a stride of N or N^2 coefficients (0.54 -> 0.53 ms per pass and box at 512^3;
// more loads in flight per wave, or one box per launch against three, measured
// the same).  One such workgroup runs on a CU at a time (dynamic LDS the
// kernel does not touch, 96 KB of the CU's 160 KB): fewer rows in flight at
// once are served faster (x / y pass over u, v, w at three waves per SIMD
// 1.62 / 1.54 ms, at two 1.57 / 1.50, at one 1.44 / 1.41).
constexpr unsigned kInplaceLdsCap = 96 * 1024;  // 1 wave per SIMD

__global__ __launch_bounds__(256) void prefilter_inplace_kernel(CoefSet cs, unsigned bpf, Geom g, int axis,
                                                                 const double* __restrict__ piv)
{
    const unsigned f = blockIdx.x / bpf;
    prefilter_inplace_one((long)(blockIdx.x - f * bpf) * blockDim.x + threadIdx.x, cs.c[f], g, axis, piv);
}

static unsigned blocks_for(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// x then y pass over nf (1..3) boxes
static hipError_t prefilter_xy(const CoefSet& cs, int nf, const Geom& g, const double* pivots, hipStream_t s)
{
    const unsigned bx = blocks_for((long)g.nt * g.N, 256), by = blocks_for((long)g.N * g.N, 256);
    hipLaunchKernelGGL(prefilter_inplace_kernel, dim3(bx * nf), dim3(256), kInplaceLdsCap, s, cs, bx, g, 1, pivots);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(prefilter_inplace_kernel, dim3(by * nf), dim3(256), kInplaceLdsCap, s, cs, by, g, 2, pivots);
    return hipGetLastError();
}

static hipError_t prefilter_xy(double* coef, const Geom& g, const double* pivots, hipStream_t s)
{
    return prefilter_xy(CoefSet{{coef, coef, coef}}, 1, g, pivots, s);
}

static hipError_t launch_z(const ZArgs& z, hipStream_t s)
{
    const unsigned blocks = blocks_for((long)z.g.nt * z.g.nt, 4);
    return with_chunks(z_chunks(z.g.nt), [&](auto k) {
        hipLaunchKernelGGL(prefilter_z_kernel<decltype(k)::value>, dim3(blocks), dim3(256), 0, s, z);
        return hipGetLastError();
    });
}

hipError_t launch_prefilter(const double* W, long plane, int q, double gam, const Geom& g, const double* pivots,
                            double* coef, hipStream_t s)
{
    if (q < 0 || q > 4 || g.nt != g.nc + 6 || g.N != g.nt + 2 || !z_chunks(g.nt)) return hipErrorInvalidValue;
    hipError_t e = launch_z(ZArgs{W, plane, q, 1, gam, g, pivots, {coef, nullptr, nullptr}}, s);
    if (e != hipSuccess) return e;
    return prefilter_xy(coef, g, pivots, s);
}

hipError_t launch_prefilter_velocity(const double* W, long plane, double gam, const Geom& g, const double* pivots,
                                     double* const coef[3], hipStream_t s)
{
    if (g.nt != g.nc + 6 || g.N != g.nt + 2 || !z_chunks(g.nt)) return hipErrorInvalidValue;
    hipError_t e = launch_z(ZArgs{W, plane, 1, 3, gam, g, pivots, {coef[0], coef[1], coef[2]}}, s);
    if (e != hipSuccess) return e;
    return prefilter_xy(CoefSet{{coef[0], coef[1], coef[2]}}, 3, g, pivots, s);
}

// Laplacian of the spline c at every padded knot (eval_UBspline_3d_d_vgl at the
// cell centres, update.cpp:1218-1229).  At a knot the uniform cubic B-spline
// weights are (1, 4, 1)/6 for the value and (1, -2, 1) delta^-2 for the second
// derivative, so the Laplacian is a 27-point stencil on the coefficients.
// The result lands at the knot positions of the coefficient box `lap`.
__global__ void lap_knots_kernel(const double* __restrict__ c, Geom g, double* __restrict__ lap)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nt = g.nt, N = g.N;
    if (t >= nt * nt * nt) return;
    const int a = (int)(t / (nt * nt)), b = (int)((t / nt) % nt), k = (int)(t % nt);
    const double w[3] = {1.0 / 6.0, 4.0 / 6.0, 1.0 / 6.0}, d2[3] = {1.0, -2.0, 1.0};
    double hyy = 0, hxx = 0, hzz = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            const double* row = c + ((long)(a + i) * N + (b + j)) * N + k;
            const double rv = w[0] * row[0] + w[1] * row[1] + w[2] * row[2];
            const double r2 = d2[0] * row[0] + d2[1] * row[1] + d2[2] * row[2];
            hyy += d2[i] * w[j] * rv;
            hxx += w[i] * d2[j] * rv;
            hzz += w[i] * w[j] * r2;
        }
    lap[((long)a * N + b) * N + k] = (hyy + hxx + hzz) * (g.dinv * g.dinv);
}

hipError_t launch_lap_spline(const double* c, const Geom& g, const double* pivots, double* lap, hipStream_t s)
{
    if (c == lap || g.nt != g.nc + 6 || g.N != g.nt + 2 || !z_chunks(g.nt)) return hipErrorInvalidValue;
    const long n = (long)g.nt * g.nt * g.nt;
    hipLaunchKernelGGL(lap_knots_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, c, g, lap);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = with_chunks(z_chunks(g.nt), [&](auto k) {
        hipLaunchKernelGGL(knots_z_kernel<decltype(k)::value>, dim3(blocks_for((long)g.nt * g.nt, 4)), dim3(256), 0, s,
                           lap, g, pivots);
        return hipGetLastError();
    });
    if (e != hipSuccess) return e;
    return prefilter_xy(lap, g, pivots, s);
}

// Block order for the tracer gathers: the hardware deals workgroups
// round-robin over the 8 XCDs, each with its own L2.  Block b runs the tracers
// of block (b mod 8) G/8 + b/8 (G a multiple of 8), so each XCD walks a
// contiguous range of the block-sorted tracers (lpt_sort.hip) and neighbouring
// cell blocks, whose stencils share cache lines, meet in the same L2.  At
// 512^3 with 2^22 tracers the two together bring the bytes fetched past L2 per
// sample / Heun launch from 12.9 / 15.0 GiB to 4.5 / 4.7 GiB (the u, v, w
// coefficients are 3.1 GiB); the remap with the row-major cell sort left them
// at 12.8 / 14.9 GiB.
__device__ __forceinline__ unsigned xcd_block()
{
    const unsigned b = blockIdx.x, nb = gridDim.x;
    return (nb & 7) ? b : (b & 7) * (nb >> 3) + (b >> 3);
}

// Resident tracer waves per SIMD for the gathers at 512^3, 2^22 tracers:
// more waves in flight fetch more distinct stencils at once and lose cache
// hits (sample at 64 VGPRs, up to 8 waves: 1.90 ms; capped at 4: 1.81 ms, at
// 3, 5, 6: 1.86, 1.84, 1.87 ms; Heun at 6 waves: 1.48 ms, capped at 3: 1.40
// ms, at 2 and 4: 1.47, 1.45 ms).  The cap is dynamic LDS that the kernels do
// not touch: 160 KB of LDS per CU hold four workgroups of 36 KB or three of
// 48 KB, a workgroup being one wave per SIMD.
constexpr unsigned kSampleLdsCap = 36 * 1024;  // 4 waves per SIMD
constexpr unsigned kHeunLdsCap = 48 * 1024;    // 3 waves per SIMD

__global__ __launch_bounds__(256) void sample_kernel(long np, const double* __restrict__ pos, const double* __restrict__ cu,
                                                     const double* __restrict__ cv, const double* __restrict__ cw, Geom g,
                                                     double* __restrict__ U, double* __restrict__ G)
{
    const long t = (long)xcd_block() * blockDim.x + threadIdx.x;  // four threads per tracer
    if ((t >> 2) < np) sample_quad(t >> 2, (int)(t & 3), pos, cu, cv, cw, g, U, G);
}

hipError_t launch_sample(long np, const double* pos, const double* const cu[3], const Geom& g, double* U, double* G,
                         hipStream_t s)
{
    if (np <= 0) return hipSuccess;
    hipLaunchKernelGGL(sample_kernel, dim3(blocks_for(4 * np, 256)), dim3(256), kSampleLdsCap, s, np, pos, cu[0], cu[1],
                       cu[2], g, U, G);
    return hipGetLastError();
}

// Particle::getParticleProp (ParticleTrack.cpp:99-113): value + gradient of rho,
// u, v, w and of the three velocity Laplacians, value + gradient + Hessian of p,
// at each particle.  cs = {rho, u, v, w, p, lap u, lap v, lap w} splines.  Row
// tag[p] of out receives the kPropWidth members in the order of the reference's
// Particle (ParticleTrack.h:21-22): DEN, U[3], P, gradD[3], gradU[3], gradV[3],
// gradW[3], gradP[3], lapU, lapV, lapW, hessP[9], gradLapU[3], gradLapV[3],
// gradLapW[3]; vectors over (y, x, z), hessP row-major over (y, x, z).
__global__ __launch_bounds__(256) void props_kernel(long np, const double* __restrict__ pos, const int* __restrict__ tag,
                                                    const double* const* __restrict__ cs, Geom g,
                                                    double* __restrict__ out)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    if (g.drop && isnan(pos[3 * p])) {  // dropped tracer: a row of NaN
        for (int k = 0; k < kPropWidth; k++) out[(long)tag[p] * kPropWidth + k] = __builtin_nan("");
        return;
    }
    const Loc Ly = locate(pos[3 * p + 0], g, true);
    const Loc Lx = locate(pos[3 * p + 1], g, true);
    const Loc Lz = locate(pos[3 * p + 2], g, true);
    // second-derivative weights of the cubic B-spline
    double hy[4], hx[4], hz[4];
    {
        const double d2 = g.dinv * g.dinv;
        const Loc* L3[3] = {&Ly, &Lx, &Lz};
        double* H3[3] = {hy, hx, hz};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const double t = L3[a]->t;
            H3[a][0] = (1 - t) * d2;
            H3[a][1] = (3 * t - 2) * d2;
            H3[a][2] = (1 - 3 * t) * d2;
            H3[a][3] = t * d2;
        }
    }
    const long N = g.N;
    double* o = out + (long)tag[p] * kPropWidth;
    // output slots of value and gradient per field
    const int vslot[8] = {0, 1, 2, 3, 4, 20, 21, 22};
    const int gslot[8] = {5, 8, 11, 14, 17, 32, 35, 38};
    for (int f = 0; f < 8; f++) {
        double v = 0, gy = 0, gx = 0, gz = 0, hyy = 0, hyx = 0, hyz = 0, hxx = 0, hxz = 0, hzz = 0;
        const double* cf = cs[f];
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) {
                const double* row = cf + ((Ly.i + a) * N + (Lx.i + b)) * N + Lz.i;
                double rv = 0, rd = 0, rh = 0;
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    rv += Lz.w[c] * row[c];
                    rd += Lz.d[c] * row[c];
                    rh += hz[c] * row[c];
                }
                const double wab = Ly.w[a] * Lx.w[b];
                v += wab * rv;
                gz += wab * rd;
                gy += Ly.d[a] * Lx.w[b] * rv;
                gx += Ly.w[a] * Lx.d[b] * rv;
                if (f == 4) {
                    hyy += hy[a] * Lx.w[b] * rv;
                    hxx += Ly.w[a] * hx[b] * rv;
                    hzz += wab * rh;
                    hyx += Ly.d[a] * Lx.d[b] * rv;
                    hyz += Ly.d[a] * Lx.w[b] * rd;
                    hxz += Ly.w[a] * Lx.d[b] * rd;
                }
            }
        o[vslot[f]] = v;
        o[gslot[f] + 0] = gy;
        o[gslot[f] + 1] = gx;
        o[gslot[f] + 2] = gz;
        if (f == 4) {
            o[23] = hyy;
            o[24] = hyx;
            o[25] = hyz;
            o[26] = hyx;
            o[27] = hxx;
            o[28] = hxz;
            o[29] = hyz;
            o[30] = hxz;
            o[31] = hzz;
        }
    }
}

hipError_t launch_props(long np, const double* pos, const int* tag, const double* const* cs_dev, const Geom& g,
                        double* out, hipStream_t s)
{
    if (np <= 0) return hipSuccess;
    hipLaunchKernelGGL(props_kernel, dim3(blocks_for(np, 256)), dim3(256), 0, s, np, pos, tag, cs_dev, g, out);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void heun_kernel(long np, double* __restrict__ pos, const double* __restrict__ U,
                                                   const double* __restrict__ cu, const double* __restrict__ cv,
                                                   const double* __restrict__ cw, Geom g, const double* __restrict__ dtp)
{
    const long t = (long)xcd_block() * blockDim.x + threadIdx.x;  // four threads per tracer
    if ((t >> 2) < np) heun_quad(t >> 2, (int)(t & 3), pos, U, cu, cv, cw, g, *dtp);
}

hipError_t launch_heun(long np, double* pos, const double* U, const double* const cn[3], const Geom& g,
                       const double* dt, hipStream_t s)
{
    if (np <= 0) return hipSuccess;
    hipLaunchKernelGGL(heun_kernel, dim3(blocks_for(4 * np, 256)), dim3(256), kHeunLdsCap, s, np, pos, U, cn[0], cn[1],
                       cn[2], g, dt);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Ensemble: the prefilter passes, sampling and Heun of every member in one
// launch each.  Block b belongs to the entry e with first[e] <= b < first[e+1]
// (binary search; uniform over the block) and runs the single-context body with
// that entry's arguments and block index b - first[e].
// ---------------------------------------------------------------------------
template <int K>
__global__ __launch_bounds__(256) void ens_prefilter_z_kernel(const ZArgs* __restrict__ zs,
                                                              const int* __restrict__ first, int n)
{
    const int e = ens_entry(first, n, blockIdx.x);
    const ZArgs z = zs[e];
    const long line = (long)(blockIdx.x - first[e]) * 4 + (threadIdx.x >> 6);
    if (line >= (long)z.g.nt * z.g.nt) return;  // uniform over the wave
    prefilter_z_line<K>(line, z);
}

__global__ __launch_bounds__(64) void ens_prefilter_inplace_kernel(const EnsLine* __restrict__ lines,
                                                                   const int* __restrict__ first, int n, int axis)
{
    const int e = ens_entry(first, n, blockIdx.x);
    const EnsLine L = lines[e];
    prefilter_inplace_one((long)(blockIdx.x - first[e]) * 64 + threadIdx.x, L.coef, L.g, axis, L.piv);
}

__global__ __launch_bounds__(256) void ens_sample_kernel(const EnsPart* __restrict__ parts, const int* __restrict__ first,
                                                         int n)
{
    const int e = ens_entry(first, n, blockIdx.x);
    const EnsPart P = parts[e];
    const long t = (long)(blockIdx.x - first[e]) * 256 + threadIdx.x;  // four threads per tracer
    if ((t >> 2) < P.np) sample_quad(t >> 2, (int)(t & 3), P.pos, P.c[0], P.c[1], P.c[2], P.g, P.U, P.G);
}

__global__ __launch_bounds__(256) void ens_heun_kernel(const EnsPart* __restrict__ parts, const int* __restrict__ first,
                                                       int n)
{
    const int e = ens_entry(first, n, blockIdx.x);
    const EnsPart P = parts[e];
    const long t = (long)(blockIdx.x - first[e]) * 256 + threadIdx.x;  // four threads per tracer
    if ((t >> 2) < P.np) heun_quad(t >> 2, (int)(t & 3), P.pos, P.U, P.c[0], P.c[1], P.c[2], P.g, *P.dt);
}

int ens_z_blocks(const Geom& g) { return (int)blocks_for((long)g.nt * g.nt, 4); }

int ens_xy_blocks(const Geom& g, int axis)
{
    return (int)blocks_for(axis == 1 ? (long)g.nt * g.N : (long)g.N * g.N, 64);
}

int ens_z_chunks(int nt_max) { return z_chunks(nt_max); }

int ens_particle_blocks(long np) { return (int)blocks_for(4 * np, 256); }  // four threads per tracer

hipError_t launch_ens_prefilter(const ZArgs* zs, const int* zfirst, int nz, int zblk, int zchunks,
                                const EnsLine* lines, const int* lfirst, int nl, const int lblk[2], hipStream_t s)
{
    if (nz <= 0) return hipSuccess;
    hipError_t e = with_chunks(zchunks, [&](auto k) {
        hipLaunchKernelGGL(ens_prefilter_z_kernel<decltype(k)::value>, dim3(zblk), dim3(256), 0, s, zs, zfirst, nz);
        return hipGetLastError();
    });
    for (int axis = 1; axis <= 2 && e == hipSuccess; axis++) {
        hipLaunchKernelGGL(ens_prefilter_inplace_kernel, dim3(lblk[axis - 1]), dim3(64), 0, s, lines,
                           lfirst + (axis - 1) * (nl + 1), nl, axis);
        e = hipGetLastError();
    }
    return e;
}

hipError_t launch_ens_sample(const EnsPart* parts, const int* first, int n, int nblk, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(ens_sample_kernel, dim3(nblk), dim3(256), 0, s, parts, first, n);
    return hipGetLastError();
}

hipError_t launch_ens_heun(const EnsPart* parts, const int* first, int n, int nblk, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(ens_heun_kernel, dim3(nblk), dim3(256), 0, s, parts, first, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Statistics.  Velocity gradients at cell centres come from the spline at its
// knots, where the uniform cubic B-spline reduces to the 3-point stencils
// (1,4,1)/6 (value) and (-1,0,1)/(2 delta) (derivative) on the coefficients.
// Block partial sums (wave shuffles + LDS) go to red[]; the host adds the
// kStatBlocks partials in a fixed order.
// ---------------------------------------------------------------------------
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double* out)
{
    __shared__ double sm[4][K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        double x = v[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        v[k] = x;
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < K; k++) sm[wv][k] = v[k];
    __syncthreads();
    if (threadIdx.x < K) out[threadIdx.x] = sm[0][threadIdx.x] + sm[1][threadIdx.x] + sm[2][threadIdx.x] + sm[3][threadIdx.x];
}

__global__ __launch_bounds__(256) void stats_kernel(const double* __restrict__ W, long plane, const double* __restrict__ cu,
                                                    const double* __restrict__ cv, const double* __restrict__ cw, Geom g,
                                                    StatParams sp, double* __restrict__ red)
{
    // sums: 0 tke, 1 diss, 2 sijsij, 3 enstrophy, 4-6 d2, 7-9 d3, 10-12 d4, 13 P, 14 T, 15 1/rho
    double acc[16];
#pragma unroll
    for (int k = 0; k < 16; k++) acc[k] = 0.0;
    const int nc = g.nc;
    const long N = g.N;
    const double w0 = 1.0 / 6.0, w1 = 4.0 / 6.0;
    const double d = 0.5 * g.dinv;
    const double* cs[3] = {cu, cv, cw};
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < plane; c += (long)gridDim.x * blockDim.x) {
        const int k = (int)(c % nc), i = (int)((c / nc) % nc), j = (int)(c / ((long)nc * nc));
        const double den = W[c], ir = 1.0 / den;
        const double uu = W[plane + c] * ir, vv = W[2 * plane + c] * ir, ww = W[3 * plane + c] * ir;
        const double Vsqr = uu * uu + vv * vv + ww * ww;
        const double P = (den * (sp.gam - 1)) * (W[4 * plane + c] * ir - 0.5 * Vsqr);
        const double T = P / (sp.R * den);
        double val[3], gr[3][3];  // gr[f] = d/dy, d/dx, d/dz
#pragma unroll
        for (int f = 0; f < 3; f++) {
            // knot (j+3, i+3, k+3): coefficient block starts at (j+3, i+3, k+3) with t = 0
            const double* base = cs[f] + ((long)(j + 3) * N + (i + 3)) * N + (k + 3);
            double v = 0, gy = 0, gx = 0, gz = 0;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const double wa = a == 1 ? w1 : w0;
                const double da = a == 0 ? -d : (a == 2 ? d : 0.0);
#pragma unroll
                for (int b = 0; b < 3; b++) {
                    const double wb = b == 1 ? w1 : w0;
                    const double db = b == 0 ? -d : (b == 2 ? d : 0.0);
                    const double* row = base + (a * N + b) * N;
                    const double rv = w0 * row[0] + w1 * row[1] + w0 * row[2];
                    const double rd = d * (row[2] - row[0]);
                    v += wa * wb * rv;
                    gz += wa * wb * rd;
                    gy += da * wb * rv;
                    gx += wa * db * rv;
                }
            }
            val[f] = v;
            gr[f][0] = gy;
            gr[f][1] = gx;
            gr[f][2] = gz;
        }
        const double dudx = gr[0][1], dudy = gr[0][0], dudz = gr[0][2];
        const double dvdx = gr[1][1], dvdy = gr[1][0], dvdz = gr[1][2];
        const double dwdx = gr[2][1], dwdy = gr[2][0], dwdz = gr[2][2];
        const double x = T / sp.T0;
        const double mu = sp.mu0 * (x * sqrt(x)) * (sp.T0 + 110.4) / (T + 110.4);
        acc[0] += den * (val[0] * val[0] + val[1] * val[1] + val[2] * val[2]);
        acc[1] += mu * (dudx * dudx + dudy * dudy + dudz * dudz + dvdx * dvdx + dvdy * dvdy + dvdz * dvdz +
                        dwdx * dwdx + dwdy * dwdy + dwdz * dwdz);
        const double s1 = dudy + dvdx, s2 = dudz + dwdx, s3 = dwdy + dvdz;
        acc[2] += dudx * dudx + 0.5 * s1 * s1 + 0.5 * s2 * s2 + 0.5 * s3 * s3 + dvdy * dvdy + dwdz * dwdz;
        const double o1 = dwdy - dvdz, o2 = dudz - dwdx, o3 = dvdx - dudy;
        acc[3] += o1 * o1 + o2 * o2 + o3 * o3;
        const double dd[3] = {dudx, dvdy, dwdz};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const double s = dd[a] * dd[a];
            acc[4 + a] += s;
            acc[7 + a] += s * dd[a];
            acc[10 + a] += s * s;
        }
        acc[13] += P;
        acc[14] += T;
        acc[15] += ir;
    }
    block_sum<16>(acc, red + 32 * blockIdx.x);
}

__global__ __launch_bounds__(256) void stats2_kernel(const double* __restrict__ W, long plane, StatParams sp, double pm,
                                                     double tm, double vm, double* __restrict__ red)
{
    double acc[3] = {0, 0, 0};
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < plane; c += (long)gridDim.x * blockDim.x) {
        const double den = W[c], ir = 1.0 / den;
        const double uu = W[plane + c] * ir, vv = W[2 * plane + c] * ir, ww = W[3 * plane + c] * ir;
        const double P = (den * (sp.gam - 1)) * (W[4 * plane + c] * ir - 0.5 * (uu * uu + vv * vv + ww * ww));
        const double T = P / (sp.R * den);
        acc[0] += (pm - P) * (pm - P);
        acc[1] += (tm - T) * (tm - T);
        acc[2] += (vm - ir) * (vm - ir);
    }
    block_sum<3>(acc, red + 32 * blockIdx.x + 16);
}

hipError_t launch_stats(const double* W, long plane, const double* const cu[3], const Geom& g, const StatParams& sp,
                        double* red, hipStream_t s)
{
    hipLaunchKernelGGL(stats_kernel, dim3(kStatBlocks), dim3(256), 0, s, W, plane, cu[0], cu[1], cu[2], g, sp, red);
    return hipGetLastError();
}

hipError_t launch_stats2(const double* W, long plane, const StatParams& sp, double pm, double tm, double vm,
                         double* red, hipStream_t s)
{
    hipLaunchKernelGGL(stats2_kernel, dim3(kStatBlocks), dim3(256), 0, s, W, plane, sp, pm, tm, vm, red);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void pstats_kernel(long np, const double* __restrict__ U, const double* __restrict__ G,
                                                     double* __restrict__ red)
{
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (long)gridDim.x * blockDim.x) {
        const double* u = U + 3 * p;
        const double* gg = G + 9 * p;
        acc[0] += u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        const double dudy = gg[0], dudx = gg[1], dudz = gg[2];
        const double dvdy = gg[3], dvdx = gg[4], dvdz = gg[5];
        const double dwdy = gg[6], dwdx = gg[7], dwdz = gg[8];
        const double s1 = dudy + dvdx, s2 = dudz + dwdx, s3 = dwdy + dvdz;
        acc[1] += dudx * dudx + 0.5 * s1 * s1 + 0.5 * s2 * s2 + 0.5 * s3 * s3 + dvdy * dvdy + dwdz * dwdz;
        acc[2] += dudx * dudx;
        acc[3] += dudx * dudx * dudx;
        acc[4] += dwdz * dwdz;
        acc[5] += dwdz * dwdz * dwdz;
    }
    block_sum<6>(acc, red + 8 * blockIdx.x);
}

hipError_t launch_pstats(long np, const double* U, const double* G, double* red, hipStream_t s)
{
    hipLaunchKernelGGL(pstats_kernel, dim3(kStatBlocks), dim3(256), 0, s, np, U, G, red);
    return hipGetLastError();
}

}  // namespace lpt
