// gfx950 kernels of the post-relax reductions of one network's state:
// displacements u = x - x0 from the positions before the first strain step
// (x0) to the current ones (x), both in device node order.
//   non-affinity  (na_calc.cpp:308-326):   mean over nodes of |u - u_aff|^2,
//                 u_aff = ((y0 - miny) shear, (y0 - miny) axial)
//   bond strains  (strain_calc.cpp:96-120): (l_f - l_i) / l_i per bond, edge
//                 order, l_f from d_i + u_j - u_i
//   axial energy  (axial_contribution.cpp:173-190): k/2 sum (u_ij . rhat_ij)^2,
//                 rhat from the initial geometry
// A bond across the y boundary of a Lees-Edwards box sees the image of its
// second node, displaced by the box shift accumulated since x0.
// Partial sums come from a fixed grid (grid-stride loops, an LDS tree per
// workgroup) and are added on the host in workgroup order: the result does
// not depend on scheduling.
#include <hip/hip_runtime.h>

#include <mutex>

#include "device.hpp"

namespace nr {

// Workspace of nr_batch_analyze_noalloc, one copy per GPU for the life of the
// process: the workgroup partials (4 kAnalysisBlocks doubles), then the bond
// strains (kNoAllocMaxBonds doubles), then a staged bond list
// (kNoAllocMaxBonds int4 = 2 kNoAllocMaxBonds doubles).  Host code reaches it
// only through the address hipGetSymbolAddress gives for the current device.
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace {

template <int B>
__device__ inline double tree_sum(double v, double *s)
{
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = B / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    const double r = s[0];
    __syncthreads();
    return r;
}

template <int B>
__device__ inline double tree_max(double v, double *s)
{
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = B / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + o]);
        __syncthreads();
    }
    const double r = s[0];
    __syncthreads();
    return r;
}

template <int B>
__global__ __launch_bounds__(B) void nonaffine_kernel(const double2 *__restrict__ x,
                                                      const double2 *__restrict__ x0, int n,
                                                      double miny, double shear, double axial,
                                                      double *part)
{
    __shared__ double s[B];
    double acc = 0.0;
    for (int k = blockIdx.x * B + threadIdx.x; k < n; k += gridDim.x * B) {
        const double2 p = x[k], q = x0[k];
        const double ux = p.x - q.x, uy = p.y - q.y;
        const double ex = (q.y - miny) * shear, ey = (q.y - miny) * axial;
        acc += (ux - ex) * (ux - ex) + (uy - ey) * (uy - ey);
    }
    acc = tree_sum<B>(acc, s);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

template <int B>
__global__ __launch_bounds__(B) void bond_kernel(const double2 *__restrict__ x,
                                                 const double2 *__restrict__ x0,
                                                 const int4 *__restrict__ bonds, int ne, double W,
                                                 double H, double shift, double *strain,
                                                 double *part)
{
    __shared__ double s[B];
    double ax = 0.0, sum = 0.0, mx = 0.0;
    for (int e = blockIdx.x * B + threadIdx.x; e < ne; e += gridDim.x * B) {
        const int4 b = bonds[e];
        const double2 pi0 = x0[b.x], pj0 = x0[b.y], pi = x[b.x], pj = x[b.y];
        const double dxi = pj0.x + b.z * W - pi0.x, dyi = pj0.y + b.w * H - pi0.y;
        const double uxi = pi.x - pi0.x, uyi = pi.y - pi0.y;
        const double uxj = pj.x - pj0.x + b.w * shift, uyj = pj.y - pj0.y;
        const double dxf = dxi + uxj - uxi, dyf = dyi + uyj - uyi;
        const double li = sqrt(dxi * dxi + dyi * dyi), lf = sqrt(dxf * dxf + dyf * dyf);
        const double st = (lf - li) / li;
        if (strain) strain[e] = st;
        const double proj = (uxj - uxi) * (dxi / li) + (uyj - uyi) * (dyi / li);
        ax += 0.5 * proj * proj;
        sum += st;
        mx = fmax(mx, fabs(st));
    }
    ax = tree_sum<B>(ax, s);
    sum = tree_sum<B>(sum, s);
    mx = tree_max<B>(mx, s);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = ax;
        part[3 * blockIdx.x + 1] = sum;
        part[3 * blockIdx.x + 2] = mx;
    }
}

constexpr int kAnThreads = 256;

}  // namespace

int launch_analysis(const double2 *x, const double2 *x0, int n, double miny, double shear,
                    double axial, const int4 *bonds, int ne, double W, double H, double shift,
                    double *strain, double *part, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(nonaffine_kernel<kAnThreads>, dim3(kAnalysisBlocks), dim3(kAnThreads), 0,
                       st, x, x0, n, miny, shear, axial, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
    if (ne > 0) {
        hipLaunchKernelGGL(bond_kernel<kAnThreads>, dim3(kAnalysisBlocks), dim3(kAnThreads), 0,
                           st, x, x0, bonds, ne, W, H, shift, strain, part + kAnalysisBlocks);
        e = hipGetLastError();
    }
    return (int)e;
}

// Device address of `scratch` on the current device, looked up once per
// device (each GPU holds its own copy at its own address).
int analysis_scratch(int device, double **ptr)
{
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static double *addr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) return (int)hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        double *p = nullptr;
        const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        addr[device] = p;
    }
    *ptr = addr[device];
    return 0;
}

}  // namespace nr
