// gfx950 kernels of the small-strain linear solve (the reference's
// sparse_relax.cpp, whose SuiteSparse QR is replaced by conjugate gradients on
// the device): K_rr w = b over the free DOFs, K the assembled stiffness in
// CSR (FP64), all vectors and scalars resident.  One CG iteration k is two
// launches:
//   cg_dir_spmv  r.r of the current r (from the last update's partials), the
//                stopping rule, beta = r.r / r.r_old, then for its rows
//                p = r + beta p_old (written to the other p buffer) and
//                q = P K p, with p at each column formed on the fly from r and
//                p_old (the same value the row's owner writes), partials of p.q
//   cg_update    alpha = r.r / p.q; w += alpha p; r -= alpha q; partials of r.r
// (the direction update of the textbook iteration, fused into the SpMV, so
// that an iteration waits for two launches instead of three; the arithmetic
// is the textbook order's, value for value).  Scalars are re-reduced from
// kLinBlocks per-workgroup partials in a fixed order by every workgroup that
// needs them (no atomics; the result does not depend on scheduling).
// Iterations are captured in chunks of an even number k = 0, 1, ...: r.r and
// the done flag are kept per iteration parity (k & 1), and p alternates
// between two buffers by the same parity, so no kernel writes what another
// workgroup of the same launch reads.  Once a done flag is up every kernel
// returns at once (cg_dir_spmv carrying it forward), so replays never
// overshoot.  Scalars sc: [0, 1] r.r by parity, [2] (tol x scale)^2, [3, 4]
// done before iteration k by parity, [5] iterations before this chunk,
// [6] iterations when done, [7] the last r.r, [8] set when p.q <= 0 (no
// descent left).
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
__global__ __launch_bounds__(B) void cg_dir_spmv(int n, const int *__restrict__ rowptr,
                                                 const int *__restrict__ col,
                                                 const double *__restrict__ val,
                                                 const unsigned char *__restrict__ free_dof,
                                                 const double *__restrict__ r,
                                                 const double *__restrict__ p_old,
                                                 double *__restrict__ p_new, double *__restrict__ q,
                                                 double *part_pq, const double *part_rr, double *sc,
                                                 int k, int max_iter)
{
    __shared__ double s[B];
    const int par = k & 1;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (sc[3 + par] != 0.0 || sc[8] != 0.0) {
        if (lead) sc[3 + (par ^ 1)] = 1.0;
        return;
    }
    const double it = sc[5] + (double)k;  // iterations done
    const bool first = it == 0.0;
    // r.r of the current r: the host's for the first iteration
    const double rr1 = first ? sc[par] : sum_partials<B>(part_rr, s);
    if (!(rr1 > sc[2]) || it >= (double)max_iter) {
        if (lead) {
            sc[3 + (par ^ 1)] = 1.0;
            sc[6] = it;
            sc[7] = rr1;
        }
        return;
    }
    const double rr0 = sc[par];
    const double beta = !first && rr0 > 0.0 ? rr1 / rr0 : 0.0;
    double acc = 0.0;
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) {
        const double pi = fma(beta, p_old[i], r[i]);
        p_new[i] = pi;
        double v = 0.0;
        if (free_dof[i]) {
            const int lo = rowptr[i], hi = rowptr[i + 1];
            for (int e = lo; e < hi; e++) {
                const int j = col[e];
                v = fma(val[e], fma(beta, p_old[j], r[j]), v);
            }
        }
        q[i] = v;
        acc = fma(pi, v, acc);
    }
    acc = lin_tree_sum<B>(acc, s);
    if (threadIdx.x == 0) part_pq[blockIdx.x] = acc;
    if (lead) {
        sc[par ^ 1] = rr1;
        sc[7] = rr1;
    }
}

template <int B>
__global__ __launch_bounds__(B) void cg_update(int n, const double *__restrict__ p,
                                               const double *__restrict__ q, double *w, double *r,
                                               const double *part_pq, double *part_rr, double *sc,
                                               int k)
{
    __shared__ double s[B];
    const int par = k & 1;
    if (sc[3 + (par ^ 1)] != 0.0 || sc[8] != 0.0) return;
    const double pq = sum_partials<B>(part_pq, s);
    if (!(pq > 0.0)) {  // no descent left: the iteration does not move w
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            sc[8] = 1.0;
            sc[6] = sc[5] + (double)k;
        }
        return;
    }
    const double alpha = sc[par ^ 1] / pq;
    double acc = 0.0;
    for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) {
        w[i] = fma(alpha, p[i], w[i]);
        const double ri = fma(-alpha, q[i], r[i]);
        r[i] = ri;
        acc = fma(ri, ri, acc);
    }
    acc = lin_tree_sum<B>(acc, s);
    if (threadIdx.x == 0) part_rr[blockIdx.x] = acc;
}

// iterations before the next chunk (after the chunk's last update)
__global__ void cg_chunk_end(double *sc, int chunk)
{
    if (threadIdx.x == 0) sc[5] += (double)chunk;
}

}  // namespace

int launch_cg_iteration(int n, const int *rowptr, const int *col, const double *val,
                        const unsigned char *free_dof, double *w, double *r, double *const p[2],
                        double *q, double *part, double *sc, int k, int max_iter, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    double *part_pq = part, *part_rr = part + kLinBlocks;
    const int par = k & 1;
    hipLaunchKernelGGL(cg_dir_spmv<kLinThreads>, dim3(kLinBlocks), dim3(kLinThreads), 0, st, n,
                       rowptr, col, val, free_dof, (const double *)r, (const double *)p[par],
                       p[par ^ 1], q, part_pq, (const double *)part_rr, sc, k, max_iter);
    hipLaunchKernelGGL(cg_update<kLinThreads>, dim3(kLinBlocks), dim3(kLinThreads), 0, st, n,
                       (const double *)p[par ^ 1], (const double *)q, w, r,
                       (const double *)part_pq, part_rr, sc, k);
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
