// gfx950 kernels of the small-strain linear solve (the reference's
// sparse_relax.cpp, whose SuiteSparse QR is replaced by conjugate gradients on
// the device): K_rr w = b over the free DOFs, K the assembled stiffness in
// CSR (FP64), all vectors and scalars resident.  One CG iteration is three
// launches:
//   cg_spmv   q = P K p (rows of clamped DOFs give 0), partials of p.q
//   cg_update alpha = r.r / p.q; w += alpha p; r -= alpha q; partials of r.r
//   cg_dir    beta = r.r' / r.r; p = r + beta p; stopping rule
// Scalars are re-reduced from kLinBlocks per-workgroup partials in a fixed
// order by every workgroup that needs them (no atomics; the result does not
// depend on scheduling).  Iterations are captured in chunks of an even
// number k = 0, 1, ...: r.r and the done flag are kept per iteration parity
// (k & 1), so a kernel reads the current slot while the first workgroup of
// cg_dir writes the next one, and no kernel writes what another workgroup
// of the same launch reads.  Once the flag is up every kernel returns at
// once (cg_dir carrying the flag forward), so replays never overshoot.
// Scalars sc: [0, 1] r.r by parity, [2] (tol x scale)^2, [3, 4] done by
// parity, [5] iterations before this chunk, [6] iterations when done,
// [7] the last r.r.
#include <hip/hip_runtime.h>

#include "device.hpp"

namespace nr {
namespace {

constexpr int kLinThreads = 256;

// Workgroup sum in a fixed order: a butterfly within each wave, then the
// waves in order (two barriers).
template <int B>
__device__ inline double lin_tree_sum(double v, double *s)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
#pragma unroll
    for (int w = 0; w < B / 64; w++) r += s[w];
    __syncthreads();
    return r;
}

// sum of the kLinBlocks partials at p, in a fixed order
template <int B>
__device__ inline double sum_partials(const double *p, double *s)
{
    double v = 0.0;
    for (int i = threadIdx.x; i < kLinBlocks; i += B) v += p[i];
    return lin_tree_sum<B>(v, s);
}

template <int B>
__global__ __launch_bounds__(B) void cg_spmv(int n, const int *__restrict__ rowptr,
                                             const int *__restrict__ col,
                                             const double *__restrict__ val,
                                             const unsigned char *__restrict__ free_dof,
                                             const double *__restrict__ p, double *__restrict__ q,
                                             double *part, const double *sc, int par)
{
    __shared__ double s[B];
    if (sc && sc[3 + par] != 0.0) return;
    double acc = 0.0;
    for (int r = blockIdx.x * B + threadIdx.x; r < n; r += gridDim.x * B) {
        double v = 0.0;
        if (free_dof[r]) {
            const int lo = rowptr[r], hi = rowptr[r + 1];
            for (int k = lo; k < hi; k++) v = fma(val[k], p[col[k]], v);
        }
        q[r] = v;
        acc = fma(p[r], v, acc);
    }
    acc = lin_tree_sum<B>(acc, s);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

template <int B>
__global__ __launch_bounds__(B) void cg_update(int n, const double *__restrict__ p,
                                               const double *__restrict__ q, double *w, double *r,
                                               const double *part_pq, double *part_rr,
                                               const double *sc, int par)
{
    __shared__ double s[B];
    if (sc[3 + par] != 0.0) return;
    const double pq = sum_partials<B>(part_pq, s);
    const double alpha = pq > 0.0 ? sc[par] / pq : 0.0;
    double acc = 0.0;
    for (int k = blockIdx.x * B + threadIdx.x; k < n; k += gridDim.x * B) {
        w[k] = fma(alpha, p[k], w[k]);
        const double rk = fma(-alpha, q[k], r[k]);
        r[k] = rk;
        acc = fma(rk, rk, acc);
    }
    acc = lin_tree_sum<B>(acc, s);
    if (threadIdx.x == 0) part_rr[blockIdx.x] = acc;
}

template <int B>
__global__ __launch_bounds__(B) void cg_dir(int n, const double *__restrict__ r, double *p,
                                            const double *part_pq, const double *part_rr,
                                            double *sc, int k, int max_iter)
{
    __shared__ double s[B];
    const int par = k & 1;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (sc[3 + par] != 0.0) {
        if (lead) sc[3 + (par ^ 1)] = 1.0;
        return;
    }
    const double pq = sum_partials<B>(part_pq, s);
    const double rr1 = sum_partials<B>(part_rr, s);
    const double rr0 = sc[par];
    const double it = sc[5] + (double)(k + 1);
    if (!(pq > 0.0)) {  // no descent left: the iteration did not move w
        if (lead) {
            sc[3 + (par ^ 1)] = 1.0;
            sc[6] = it - 1.0;
            sc[7] = rr0;
        }
        return;
    }
    const double beta = rr0 > 0.0 ? rr1 / rr0 : 0.0;
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B)
        p[i] = fma(beta, p[i], r[i]);
    if (lead) {
        sc[par ^ 1] = rr1;
        sc[7] = rr1;
        if (!(rr1 > sc[2]) || it >= (double)max_iter) {
            sc[3 + (par ^ 1)] = 1.0;
            sc[6] = it;
        }
    }
}

// iterations before the next chunk (after the chunk's last cg_dir)
__global__ void cg_chunk_end(double *sc, int chunk)
{
    if (threadIdx.x == 0) sc[5] += (double)chunk;
}

}  // namespace

int launch_cg_iteration(int n, const int *rowptr, const int *col, const double *val,
                        const unsigned char *free_dof, double *w, double *r, double *p, double *q,
                        double *part, double *sc, int k, int max_iter, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    double *part_pq = part, *part_rr = part + kLinBlocks;
    const int par = k & 1;
    hipLaunchKernelGGL(cg_spmv<kLinThreads>, dim3(kLinBlocks), dim3(kLinThreads), 0, st, n,
                       rowptr, col, val, free_dof, p, q, part_pq, (const double *)sc, par);
    hipLaunchKernelGGL(cg_update<kLinThreads>, dim3(kLinBlocks), dim3(kLinThreads), 0, st, n, p,
                       q, w, r, (const double *)part_pq, part_rr, (const double *)sc, par);
    hipLaunchKernelGGL(cg_dir<kLinThreads>, dim3(kLinBlocks), dim3(kLinThreads), 0, st, n, r, p,
                       (const double *)part_pq, (const double *)part_rr, sc, k, max_iter);
    return (int)hipGetLastError();
}

int launch_cg_chunk_end(double *sc, int chunk, void *stream)
{
    hipLaunchKernelGGL(cg_chunk_end, dim3(1), dim3(64), 0, (hipStream_t)stream, sc, chunk);
    return (int)hipGetLastError();
}

int launch_spmv(int n, const int *rowptr, const int *col, const double *val,
                const unsigned char *free_dof, const double *x, double *y, double *part,
                void *stream)
{
    hipLaunchKernelGGL(cg_spmv<kLinThreads>, dim3(kLinBlocks), dim3(kLinThreads), 0,
                       (hipStream_t)stream, n, rowptr, col, val, free_dof, x, y, part,
                       (const double *)nullptr, 0);
    return (int)hipGetLastError();
}

}  // namespace nr
