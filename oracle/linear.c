/*
 * Linear (small-strain) response of a network: a CPU FP64 restatement of the
 * reference's sparse linear solver, src/sparse_relax.cpp.  TEST
 * INFRASTRUCTURE ONLY: used by tests/ and bench.py's CPU-baseline leg, never
 * by the product.
 *
 * Followed:
 *   calc_stretching_elements :286-322  ks d d^T / l^3 blocks per bond, l from
 *                                      the input geometry
 *   calc_bending_elements    :325-386  the Hessian of the linearised bending
 *                                      energy below (r = hinge -> first end)
 *   calc_shearing_elements   :390-428  gel term mu (1 - rhat_x^2), mu (1 -
 *                                      rhat_y^2) on the x and y blocks only
 *   get_lhs_rhs              :560-605  the reduced system over free DOFs:
 *                                      K_rr w = -K_rN bc
 *   affine_displacement      :527-537  bc = (s (y - ymin), c (ymin - y))
 *   add_displacement         :513-523  or only the top row: (s, -c)
 *   do_relaxation_run        :676-881  u = bc + P w, residual 2-norm
 *                                      |K_rr w + K_rN bc|, energies
 *   get_stretch_shear_energy :49-70    de = du.d / L, e = ks de^2/2 +
 *                                      mu (|du|^2 - de^2)/2
 *   get_bending_energy       :73-96    e = kb (ddu x r)^2 / (2 |r|^2),
 *                                      ddu = u_a + u_b - 2 u_h
 * The reference factorises K_rr by sparse QR (SuiteSparse, not available
 * here); this restatement solves the same consistent PSD system by conjugate
 * gradients from w = 0, whose iterates stay in the range of K_rr: the result
 * is the minimum-norm solution (floppy modes and isolated nodes get no
 * displacement), which leaves the energies unchanged.  Where the reference
 * solves in strain mode it scales the shear by the x extent and then divides
 * it by the y extent (:731-739, :785); the caller passes the strains applied
 * to the field.  Lees-Edwards networks (an extension, as elsewhere in this
 * oracle) take the same form with every node free and the image of a node
 * across the y boundary displaced by (s H, -c H).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

void or_csr_free(or_csr *K)
{
    free(K->rowptr);
    free(K->col);
    free(K->val);
    memset(K, 0, sizeof(*K));
}

typedef struct {
    int r, c;
    double v;
} or_trip;

static int or_trip_cmp(const void *a, const void *b)
{
    const or_trip *x = (const or_trip *)a, *y = (const or_trip *)b;
    if (x->r != y->r) return x->r < y->r ? -1 : 1;
    if (x->c != y->c) return x->c < y->c ? -1 : 1;
    return 0;
}

/* geometry of bond b / triple arm from the input positions */
static void or_lin_bond(const or_net *net, int b, double *dx, double *dy)
{
    const int i = net->bond[2 * b], j = net->bond[2 * b + 1];
    const int wx = net->bwrap[2 * b], wy = net->bwrap[2 * b + 1];
    *dx = net->pos[2 * j] + net->width * wx + net->le_shift * wy - net->pos[2 * i];
    *dy = net->pos[2 * j + 1] + net->height * wy - net->pos[2 * i + 1];
}

static void or_lin_arm(const or_net *net, int t, double *dx, double *dy)
{
    const int h = net->tri[3 * t], a = net->tri[3 * t + 1];
    const int wx = net->twrap[4 * t], wy = net->twrap[4 * t + 1];
    *dx = net->pos[2 * a] + net->width * wx + net->le_shift * wy - net->pos[2 * h];
    *dy = net->pos[2 * a + 1] + net->height * wy - net->pos[2 * h + 1];
}

static double or_lin_kb(const or_net *net, double kb, int t)
{
    return net->tkb ? net->tkb[t] : kb;
}

/* adds v at (r, c) and, off the diagonal, at (c, r) */
static void or_put(or_trip *T, size_t *m, int r, int c, double v)
{
    T[(*m)++] = (or_trip){r, c, v};
    if (r != c) T[(*m)++] = (or_trip){c, r, v};
}

int or_linear_stiffness(const or_net *net, double ks, double kb, double mu, or_csr *K)
{
    const int n2 = 2 * net->n;
    /* entries per bond: 16 stretching + 8 gel; per triple: 6 + 2 x 15 */
    const size_t cap = (size_t)net->e * 24 + (size_t)net->t * 36 + 1;
    or_trip *T = (or_trip *)malloc(cap * sizeof(or_trip));
    size_t m = 0;
    int b, t;
    if (!T) return OR_E_ALLOC;
    memset(K, 0, sizeof(*K));
    /* stretching :286-322 */
    for (b = 0; ks > 0 && b < net->e; b++) {
        const int xi = 2 * net->bond[2 * b], xj = 2 * net->bond[2 * b + 1];
        double dx, dy;
        or_lin_bond(net, b, &dx, &dy);
        const double l = sqrt(dx * dx + dy * dy);
        const double s = ks / (l * l * l);
        const double xx = s * dx * dx, xy = s * dx * dy, yy = s * dy * dy;
        or_put(T, &m, xi, xi, xx);
        or_put(T, &m, xj, xj, xx);
        or_put(T, &m, xi, xj, -xx);
        or_put(T, &m, xi + 1, xi + 1, yy);
        or_put(T, &m, xj + 1, xj + 1, yy);
        or_put(T, &m, xi + 1, xj + 1, -yy);
        or_put(T, &m, xi, xi + 1, xy);
        or_put(T, &m, xj, xj + 1, xy);
        or_put(T, &m, xi, xj + 1, -xy);
        or_put(T, &m, xi + 1, xj, -xy);
    }
    /* bending :325-386: kb/|r|^2 g g^T with g = d(ddu x r)/du */
    for (t = 0; t < net->t; t++) {
        const double k = or_lin_kb(net, kb, t);
        if (!(k > 0)) continue;
        const int h = 2 * net->tri[3 * t], a = 2 * net->tri[3 * t + 1], c = 2 * net->tri[3 * t + 2];
        double dx, dy;
        or_lin_arm(net, t, &dx, &dy);
        const double l2 = dx * dx + dy * dy;
        const double xx = k * dx * dx / l2, yy = k * dy * dy / l2, xy = k * dx * dy / l2;
        /* DOFs and their coefficients in units of (dy, dx) */
        const int dof[6] = {h, h + 1, a, a + 1, c, c + 1};
        const double gx[6] = {-2.0, 0.0, 1.0, 0.0, 1.0, 0.0};  /* times dy */
        const double gy[6] = {0.0, 2.0, 0.0, -1.0, 0.0, -1.0}; /* times dx */
        for (int p = 0; p < 6; p++)
            for (int q = p; q < 6; q++) {
                /* (gx_p dy + gy_p dx)(gx_q dy + gy_q dx) kb / |r|^2 */
                const double v = gx[p] * gx[q] * yy + (gx[p] * gy[q] + gy[p] * gx[q]) * xy +
                                 gy[p] * gy[q] * xx;
                if (v != 0.0) or_put(T, &m, dof[p], dof[q], v);
            }
    }
    /* gel shearing :390-428 */
    for (b = 0; mu > 0 && b < net->e; b++) {
        const int xi = 2 * net->bond[2 * b], xj = 2 * net->bond[2 * b + 1];
        double dx, dy;
        or_lin_bond(net, b, &dx, &dy);
        const double inv = 1.0 / net->rest[b] / net->rest[b];
        const double mx = mu * (1.0 - dx * dx * inv), my = mu * (1.0 - dy * dy * inv);
        or_put(T, &m, xi, xi, mx);
        or_put(T, &m, xj, xj, mx);
        or_put(T, &m, xi, xj, -mx);
        or_put(T, &m, xi + 1, xi + 1, my);
        or_put(T, &m, xj + 1, xj + 1, my);
        or_put(T, &m, xi + 1, xj + 1, -my);
    }
    qsort(T, m, sizeof(or_trip), or_trip_cmp);
    K->n = n2;
    K->rowptr = (int *)calloc((size_t)n2 + 1, sizeof(int));
    K->col = (int *)malloc((m + 1) * sizeof(int));
    K->val = (double *)malloc((m + 1) * sizeof(double));
    if (!K->rowptr || !K->col || !K->val) {
        free(T);
        or_csr_free(K);
        return OR_E_ALLOC;
    }
    size_t nnz = 0;
    for (size_t q = 0; q < m; q++) {
        if (nnz > 0 && K->col[nnz - 1] == T[q].c && q > 0 && T[q - 1].r == T[q].r) {
            K->val[nnz - 1] += T[q].v;
        } else {
            K->col[nnz] = T[q].c;
            K->val[nnz] = T[q].v;
            K->rowptr[T[q].r + 1]++;
            nnz++;
        }
    }
    for (int r = 0; r < n2; r++) K->rowptr[r + 1] += K->rowptr[r];
    K->nnz = (long)nnz;
    free(T);
    return OR_OK;
}

/* y = K x */
static void or_spmv(const or_csr *K, const double *x, double *y)
{
    int r;
#pragma omp parallel for schedule(static)
    for (r = 0; r < K->n; r++) {
        double s = 0.0;
        for (int q = K->rowptr[r]; q < K->rowptr[r + 1]; q++) s = fma(K->val[q], x[K->col[q]], s);
        y[r] = s;
    }
}

/* dot product in a fixed order (64 blocks, summed in block order), so that
 * the result does not depend on the thread count */
static double or_dot(const double *a, const double *b, int n)
{
    double part[64];
    int k;
#pragma omp parallel for schedule(static)
    for (k = 0; k < 64; k++) {
        const long lo = (long)n * k / 64, hi = (long)n * (k + 1) / 64;
        double s = 0.0;
        for (long i = lo; i < hi; i++) s = fma(a[i], b[i], s);
        part[k] = s;
    }
    double s = 0.0;
    for (k = 0; k < 64; k++) s += part[k];
    return s;
}

/* load field: bc at node i, and the extra displacement of an image across
 * the y boundary */
static void or_lin_bc(const or_net *net, double s, double c, int affine, int i, double *bx,
                      double *by)
{
    const double y = net->pos[2 * i + 1];
    if (affine) {
        *bx = s * (y - net->miny);
        *by = c * (net->miny - y);
    } else {
        const int top = net->height <= 0 && y > net->maxy - 1e-10;
        *bx = top ? s : 0.0;
        *by = top ? -c : 0.0;
    }
}

void or_linear_energies(const or_net *net, double ks, double kb, double mu, const double *u,
                        double s, double c, int affine, double *e_ss, double *e_bend)
{
    /* image displacement per unit y wrap (affine field only) */
    const double ix = affine ? s * net->height : 0.0, iy = affine ? -c * net->height : 0.0;
    double es = 0.0, eb = 0.0;
    int b, t;
    for (b = 0; b < net->e; b++) {
        const int i = net->bond[2 * b], j = net->bond[2 * b + 1];
        const int wy = net->bwrap[2 * b + 1];
        double dx, dy;
        or_lin_bond(net, b, &dx, &dy);
        const double dux = u[2 * j] + ix * wy - u[2 * i];
        const double duy = u[2 * j + 1] + iy * wy - u[2 * i + 1];
        const double de = (dux * dx + duy * dy) / net->rest[b];
        es += 0.5 * ks * de * de;
        es += 0.5 * mu * (dux * dux + duy * duy - de * de);
    }
    for (t = 0; t < net->t; t++) {
        const int h = net->tri[3 * t], a = net->tri[3 * t + 1], cc = net->tri[3 * t + 2];
        const int wya = net->twrap[4 * t + 1], wyb = net->twrap[4 * t + 3];
        double dx, dy;
        or_lin_arm(net, t, &dx, &dy);
        const double l2 = dx * dx + dy * dy;
        const double dux = u[2 * a] + ix * wya + u[2 * cc] + ix * wyb - 2 * u[2 * h];
        const double duy = u[2 * a + 1] + iy * wya + u[2 * cc + 1] + iy * wyb - 2 * u[2 * h + 1];
        const double cr = dux * dy - duy * dx;
        eb += 0.5 * or_lin_kb(net, kb, t) * cr * cr / l2;
    }
    *e_ss = es;
    *e_bend = eb;
}

int or_linear_response(const or_net *net, double ks, double kb, double mu, double s, double c,
                       int affine, double tol, int max_iter, double *u, double rep[4])
{
    const int n2 = 2 * net->n;
    or_csr K;
    if (!affine && net->height > 0) return OR_E_ARG; /* no top row in a periodic box */
    int rc = or_linear_stiffness(net, ks, kb, mu, &K);
    if (rc) return rc;
    double *bc = (double *)calloc((size_t)n2, sizeof(double));
    double *g = (double *)calloc((size_t)n2, sizeof(double));
    double *w = (double *)calloc((size_t)n2, sizeof(double));
    double *r = (double *)calloc((size_t)n2, sizeof(double));
    double *p = (double *)calloc((size_t)n2, sizeof(double));
    double *q = (double *)calloc((size_t)n2, sizeof(double));
    unsigned char *fr = (unsigned char *)calloc((size_t)n2, 1);
    if (!bc || !g || !w || !r || !p || !q || !fr) {
        rc = OR_E_ALLOC;
        goto done;
    }
    for (int i = 0; i < net->n; i++) {
        or_lin_bc(net, s, c, affine, i, &bc[2 * i], &bc[2 * i + 1]);
        fr[2 * i] = (net->fcode[i] & 1) != 0;
        fr[2 * i + 1] = (net->fcode[i] & 2) != 0;
    }
    /* g = K bc, plus the images' extra displacement in Lees-Edwards mode:
     * the gradient of the energy at w = 0 */
    or_spmv(&K, bc, g);
    if (affine && net->height > 0) {
        /* each element's load is its bc difference plus the image term; the
         * image term's force is K_element applied to the image shift */
        const double ix = s * net->height, iy = -c * net->height;
        for (int b = 0; b < net->e; b++) {
            const int wy = net->bwrap[2 * b + 1];
            if (!wy) continue;
            const int i = net->bond[2 * b], j = net->bond[2 * b + 1];
            double dx, dy;
            or_lin_bond(net, b, &dx, &dy);
            const double l = sqrt(dx * dx + dy * dy);
            const double sk = ks / (l * l * l) * (dx * ix * wy + dy * iy * wy);
            g[2 * j] += sk * dx;
            g[2 * j + 1] += sk * dy;
            g[2 * i] -= sk * dx;
            g[2 * i + 1] -= sk * dy;
            if (mu > 0) {
                const double inv = 1.0 / net->rest[b] / net->rest[b];
                const double mx = mu * (1.0 - dx * dx * inv) * ix * wy;
                const double my = mu * (1.0 - dy * dy * inv) * iy * wy;
                g[2 * j] += mx;
                g[2 * j + 1] += my;
                g[2 * i] -= mx;
                g[2 * i + 1] -= my;
            }
        }
        for (int t = 0; t < net->t; t++) {
            const int wya = net->twrap[4 * t + 1], wyb = net->twrap[4 * t + 3];
            if (!wya && !wyb) continue;
            const double k = or_lin_kb(net, kb, t);
            const int h = net->tri[3 * t], a = net->tri[3 * t + 1], cc = net->tri[3 * t + 2];
            double dx, dy;
            or_lin_arm(net, t, &dx, &dy);
            const double l2 = dx * dx + dy * dy;
            const double cr = (ix * (wya + wyb)) * dy - (iy * (wya + wyb)) * dx;
            const double f = k * cr / l2;
            g[2 * h] += -2.0 * dy * f;
            g[2 * h + 1] += 2.0 * dx * f;
            g[2 * a] += dy * f;
            g[2 * a + 1] += -dx * f;
            g[2 * cc] += dy * f;
            g[2 * cc + 1] += -dx * f;
        }
    }
    /* CG on K_rr w = -g_r from w = 0, to tol x max(|g|, |K bc|): the force
     * scale of the imposed field (at an exact equilibrium g_r is round-off,
     * which must not be iterated on) */
    or_spmv(&K, bc, q);
    const double scale = fmax(sqrt(or_dot(g, g, n2)), sqrt(or_dot(q, q, n2)));
    for (int k = 0; k < n2; k++) r[k] = fr[k] ? -g[k] : 0.0;
    if (net->height > 0) {
        /* every node free in a periodic box: the two translations are null
         * vectors of K; the load is orthogonal to them up to round-off */
        double mx = 0.0, my = 0.0;
        for (int i = 0; i < net->n; i++) {
            mx += r[2 * i];
            my += r[2 * i + 1];
        }
        mx /= net->n;
        my /= net->n;
        for (int i = 0; i < net->n; i++) {
            r[2 * i] -= mx;
            r[2 * i + 1] -= my;
        }
    }
    memcpy(p, r, (size_t)n2 * sizeof(double));
    double rr = or_dot(r, r, n2);
    int it = 0;
    while (it < max_iter && sqrt(rr) > tol * scale && rr > 0.0) {
        or_spmv(&K, p, q);
        for (int k = 0; k < n2; k++)
            if (!fr[k]) q[k] = 0.0;
        const double pq = or_dot(p, q, n2);
        if (!(pq > 0.0)) break;
        const double alpha = rr / pq;
        for (int k = 0; k < n2; k++) {
            w[k] = fma(alpha, p[k], w[k]);
            r[k] = fma(-alpha, q[k], r[k]);
        }
        const double rr1 = or_dot(r, r, n2);
        const double beta = rr1 / rr;
        for (int k = 0; k < n2; k++) p[k] = fma(beta, p[k], r[k]);
        rr = rr1;
        it++;
    }
    /* true residual |K_rr w + g_r| */
    or_spmv(&K, w, q);
    double res = 0.0;
    for (int k = 0; k < n2; k++)
        if (fr[k]) res += (q[k] + g[k]) * (q[k] + g[k]);
    for (int k = 0; k < n2; k++) u[k] = bc[k] + w[k];
    or_linear_energies(net, ks, kb, mu, u, s, c, affine, &rep[0], &rep[1]);
    rep[2] = sqrt(res);
    rep[3] = it;
done:
    free(bc);
    free(g);
    free(w);
    free(r);
    free(p);
    free(q);
    free(fr);
    or_csr_free(&K);
    return rc;
}
