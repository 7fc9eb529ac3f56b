// gfx950 kernels of the small-strain linear solve (the reference's
// sparse_relax.cpp, whose SuiteSparse QR is replaced by conjugate gradients on
// the device): K_rr w = b over the free DOFs, K the assembled stiffness in
// CSR (FP64), all vectors and scalars resident.  One CG iteration is four
// launches:
//   cg_spmv   q = P K p (rows of clamped DOFs give 0), partials of p.q
//   cg_update alpha = rr / pq; w += alpha p; r -= alpha q; partials of r.r
//   cg_dir    beta = rr' / rr; p = r + beta p
//   cg_commit rr <- rr', stopping rule
// Scalars are re-reduced from kLinBlocks per-workgroup partials in a fixed
// order by every workgroup that needs them (no atomics; the result does not
// depend on scheduling).  A fourth launch (cg_commit) makes the new r.r
// current, counts the iteration and raises the done flag on the stopping
// rule (|r| <= tol x scale, r.r = 0, p.q <= 0, or the iteration cap); every
// kernel returns at once while the flag is up, so the host can queue
// iterations ahead (a captured graph of several) without overshooting.
// Scalars sc: [0] r.r, [1] p.q, [2] new r.r, [3] done, [4] iterations,
// [5] (tol x scale)^2.
#include <hip/hip_runtime.h>

#include "device.hpp"

namespace nr {
namespace {

constexpr int kLinThreads = 256;

template <int B>
__device__ inline double lin_tree_sum(double v, double *s)
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

// sum of the kLinBlocks partials at p, in a fixed order (one LDS tree)
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
                                             double *part, const double *sc)
{
    __shared__ double s[B];
    if (sc && sc[3] != 0.0) return;
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

// sc[0] = rr of the current residual, sc[1] = p.q (written by workgroup 0)
template <int B>
__global__ __launch_bounds__(B) void cg_update(int n, const double *__restrict__ p,
                                               const double *__restrict__ q, double *w, double *r,
                                               const double *part_pq, double *part_rr, double *sc)
{
    __shared__ double s[B];
    if (sc[3] != 0.0) return;
    const double pq = sum_partials<B>(part_pq, s);
    const double alpha = pq > 0.0 ? sc[0] / pq : 0.0;
    double acc = 0.0;
    for (int k = blockIdx.x * B + threadIdx.x; k < n; k += gridDim.x * B) {
        w[k] = fma(alpha, p[k], w[k]);
        const double rk = fma(-alpha, q[k], r[k]);
        r[k] = rk;
        acc = fma(rk, rk, acc);
    }
    acc = lin_tree_sum<B>(acc, s);
    if (threadIdx.x == 0) {
        part_rr[blockIdx.x] = acc;
        if (blockIdx.x == 0) sc[1] = pq;
    }
}

// new r.r into sc[2] (workgroup 0); every workgroup reads the old one in
// sc[0], which cg_commit replaces after this launch
template <int B>
__global__ __launch_bounds__(B) void cg_dir(int n, const double *__restrict__ r, double *p,
                                            const double *part_rr, double *sc)
{
    __shared__ double s[B];
    if (sc[3] != 0.0) return;
    const double rr1 = sum_partials<B>(part_rr, s);
    const double rr0 = sc[0];
    const double beta = rr0 > 0.0 ? rr1 / rr0 : 0.0;
    for (int k = blockIdx.x * B + threadIdx.x; k < n; k += gridDim.x * B)
        p[k] = fma(beta, p[k], r[k]);
    if (blockIdx.x == 0 && threadIdx.x == 0) sc[2] = rr1;
}

// sc[0] <- sc[2]: the new r.r becomes current once every cg_dir workgroup
// has read the old one (a launch of its own, stream-ordered after cg_dir)
__global__ void cg_commit(double *sc, int max_iter)
{
    if (threadIdx.x != 0 || sc[3] != 0.0) return;
    if (!(sc[1] > 0.0)) {  // no descent left: the iteration did not move w
        sc[3] = 1.0;
        return;
    }
    sc[0] = sc[2];
    sc[4] += 1.0;
    if (!(sc[0] > sc[5]) || sc[4] >= (double)max_iter) sc[3] = 1.0;
}

}  // namespace

int launch_cg_iteration(int n, const int *rowptr, const int *col, const double *val,
                        const unsigned char *free_dof, double *w, double *r, double *p, double *q,
                        double *part, double *sc, int max_iter, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    double *part_pq = part, *part_rr = part + kLinBlocks;
    hipLaunchKernelGGL(cg_spmv<kLinThreads>, dim3(kLinBlocks), dim3(kLinThreads), 0, st, n,
                       rowptr, col, val, free_dof, p, q, part_pq, (const double *)sc);
    hipLaunchKernelGGL(cg_update<kLinThreads>, dim3(kLinBlocks), dim3(kLinThreads), 0, st, n, p,
                       q, w, r, part_pq, part_rr, sc);
    hipLaunchKernelGGL(cg_dir<kLinThreads>, dim3(kLinBlocks), dim3(kLinThreads), 0, st, n, r, p,
                       part_rr, sc);
    hipLaunchKernelGGL(cg_commit, dim3(1), dim3(64), 0, st, sc, max_iter);
    return (int)hipGetLastError();
}

int launch_spmv(int n, const int *rowptr, const int *col, const double *val,
                const unsigned char *free_dof, const double *x, double *y, double *part,
                void *stream)
{
    hipLaunchKernelGGL(cg_spmv<kLinThreads>, dim3(kLinBlocks), dim3(kLinThreads), 0,
                       (hipStream_t)stream, n, rowptr, col, val, free_dof, x, y, part,
                       (const double *)nullptr);
    return (int)hipGetLastError();
}

}  // namespace nr
