// gfx950 kernels of the small-strain linear solve (the reference's
// sparse_relax.cpp, whose SuiteSparse QR is replaced by conjugate gradients on
// the device): K_rr w = b over the free DOFs, K the assembled stiffness in
// CSR (FP64), all vectors and scalars resident.  A launch serves a batch of
// independent systems (segments: networks, load steps), each with its own
// workgroups, partial sums, scalars and stopping, so that a batch pays the
// launch latency of an iteration once.  One CG iteration k is two launches:
//   cg_dir_spmv  r.r of the current r (from the last update's partials), the
//                stopping rule, beta = r.r / r.r_old, then for its rows
//                p = r + beta p_old (written to the other p buffer) and
//                q = P K p, with p at each column formed on the fly from r and
//                p_old (the same value the row's owner writes), partials of p.q
//   cg_update    alpha = r.r / p.q; w += alpha p; r -= alpha q; partials of r.r
// (the direction update of the textbook iteration, fused into the SpMV, so
// that an iteration waits for two launches instead of three; the arithmetic
// is the textbook order's, value for value).  A segment's scalars are
// re-reduced from its workgroups' partials in a fixed order by every one of
// its workgroups that needs them (no atomics; the result depends neither on
// scheduling nor on the other segments of the batch).  Iterations are
// captured in chunks of an even number k = 0, 1, ...: r.r and the done flag
// are kept per iteration parity (k & 1), and p alternates between two
// buffers by the same parity, so no kernel writes what another workgroup of
// the same launch reads.  Once a segment's done flag is up its workgroups
// return at once (cg_dir_spmv carrying the flag forward), so replays never
// overshoot.  Scalars of a segment (kLinScalars): [0, 1] r.r by parity,
// [2] (tol x scale)^2, [3, 4] done before iteration k by parity, [5]
// iterations before this chunk, [6] iterations when done, [7] the last r.r,
// [8] set when p.q <= 0 (no descent left).
#include <hip/hip_runtime.h>

#include <algorithm>

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

// sum of the count partials at p, in a fixed order
template <int B>
__device__ inline double sum_partials(const double *p, int count, double *s)
{
    double v = 0.0;
    for (int i = threadIdx.x; i < count; i += B) v += p[i];
    return lin_tree_sum<B>(v, s);
}

template <int B>
__global__ __launch_bounds__(B) void cg_dir_spmv(const LinSeg *__restrict__ segs,
                                                 const int *__restrict__ seg_of,
                                                 const int *__restrict__ rowptr,
                                                 const int *__restrict__ col,
                                                 const double *__restrict__ val,
                                                 const unsigned char *__restrict__ free_dof,
                                                 const double *__restrict__ r,
                                                 const double *__restrict__ p_old,
                                                 double *__restrict__ p_new, double *__restrict__ q,
                                                 double *part_pq, const double *part_rr,
                                                 double *sc_all, int k)
{
    __shared__ double s[B];
    const int sid = seg_of[blockIdx.x];
    const LinSeg S = segs[sid];
    double *sc = sc_all + (size_t)kLinScalars * sid;
    const int lb = blockIdx.x - S.blk0;
    const int par = k & 1;
    const bool lead = lb == 0 && threadIdx.x == 0;
    if (sc[3 + par] != 0.0 || sc[8] != 0.0) {
        if (lead) sc[3 + (par ^ 1)] = 1.0;
        return;
    }
    const double it = sc[5] + (double)k;  // iterations done
    const bool first = it == 0.0;
    // r.r of the current r: the host's for the first iteration
    const double rr1 = first ? sc[par] : sum_partials<B>(part_rr + S.blk0, S.nblk, s);
    if (!(rr1 > sc[2]) || it >= (double)S.max_iter) {
        if (lead) {
            sc[3 + (par ^ 1)] = 1.0;
            sc[6] = it;
            sc[7] = rr1;
        }
        return;
    }
    const double rr0 = sc[par];
    const double beta = !first && rr0 > 0.0 ? rr1 / rr0 : 0.0;
    const int *__restrict__ rp = rowptr + S.rp0;
    const int *__restrict__ cl = col + S.nz0;
    const double *__restrict__ vl = val + S.nz0;
    const unsigned char *__restrict__ fr = free_dof + S.fr0;
    const double *__restrict__ rs = r + S.row0;
    const double *__restrict__ po = p_old + S.row0;
    double acc = 0.0;
    for (int i = lb * B + threadIdx.x; i < S.n; i += S.nblk * B) {
        const double pi = fma(beta, po[i], rs[i]);
        p_new[S.row0 + i] = pi;
        double v = 0.0;
        if (fr[i]) {
            const int lo = rp[i], hi = rp[i + 1];
            for (int e = lo; e < hi; e++) {
                const int j = cl[e];
                v = fma(vl[e], fma(beta, po[j], rs[j]), v);
            }
        }
        q[S.row0 + i] = v;
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
__global__ __launch_bounds__(B) void cg_update(const LinSeg *__restrict__ segs,
                                               const int *__restrict__ seg_of,
                                               const double *__restrict__ p,
                                               const double *__restrict__ q, double *w, double *r,
                                               const double *part_pq, double *part_rr,
                                               double *sc_all, int k)
{
    __shared__ double s[B];
    const int sid = seg_of[blockIdx.x];
    const LinSeg S = segs[sid];
    double *sc = sc_all + (size_t)kLinScalars * sid;
    const int lb = blockIdx.x - S.blk0;
    const int par = k & 1;
    if (sc[3 + (par ^ 1)] != 0.0 || sc[8] != 0.0) return;
    const double pq = sum_partials<B>(part_pq + S.blk0, S.nblk, s);
    if (!(pq > 0.0)) {  // no descent left: the iteration does not move w
        if (lb == 0 && threadIdx.x == 0) {
            sc[8] = 1.0;
            sc[6] = sc[5] + (double)k;
        }
        return;
    }
    const double alpha = sc[par ^ 1] / pq;
    double acc = 0.0;
    for (int i = S.row0 + lb * B + threadIdx.x; i < S.row0 + S.n; i += S.nblk * B) {
        w[i] = fma(alpha, p[i], w[i]);
        const double ri = fma(-alpha, q[i], r[i]);
        r[i] = ri;
        acc = fma(ri, ri, acc);
    }
    acc = lin_tree_sum<B>(acc, s);
    if (threadIdx.x == 0) part_rr[blockIdx.x] = acc;
}

// iterations before the next chunk (after the chunk's last update), per segment
__global__ void cg_chunk_end(double *sc, int count, int chunk)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) sc[(size_t)kLinScalars * i + 5] += (double)chunk;
}

}  // namespace

int lin_segment_blocks(int n)
{
    return std::max(1, std::min(kLinBlocks, (n + kLinThreads - 1) / kLinThreads));
}

int launch_cg_iteration(const LinSeg *segs, const int *seg_of, int blocks, const int *rowptr,
                        const int *col, const double *val, const unsigned char *free_dof,
                        double *w, double *r, double *const p[2], double *q, double *part,
                        double *sc, int k, void *stream)
{
    if (blocks <= 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    double *part_pq = part, *part_rr = part + blocks;
    const int par = k & 1;
    hipLaunchKernelGGL(cg_dir_spmv<kLinThreads>, dim3(blocks), dim3(kLinThreads), 0, st, segs,
                       seg_of, rowptr, col, val, free_dof, (const double *)r,
                       (const double *)p[par], p[par ^ 1], q, part_pq, (const double *)part_rr,
                       sc, k);
    hipLaunchKernelGGL(cg_update<kLinThreads>, dim3(blocks), dim3(kLinThreads), 0, st, segs,
                       seg_of, (const double *)p[par ^ 1], (const double *)q, w, r,
                       (const double *)part_pq, part_rr, sc, k);
    return (int)hipGetLastError();
}

int launch_cg_chunk_end(double *sc, int count, int chunk, void *stream)
{
    if (count <= 0) return 0;
    hipLaunchKernelGGL(cg_chunk_end, dim3((count + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       sc, count, chunk);
    return (int)hipGetLastError();
}

}  // namespace nr
