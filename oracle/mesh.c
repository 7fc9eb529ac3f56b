/*
 * Network + FEM-gel composite (src/network_mesh_relax.cu), CPU FP64 oracle.
 * TEST INFRASTRUCTURE ONLY: used by tests/ and bench.py's CPU-baseline leg
 * as the checker, never by the product.
 *
 * A composite is a bond/bend network living on the points of a triangular
 * mesh whose facets are a compressible neo-Hookean gel.  The mesh is periodic
 * in x through duplicated points: each point on the left edge has a twin one
 * box width to its right, and the two receive the sum of their forces.
 * Functions follow the reference file:
 *   read_network_mesh :72-230 (twin pairs :219-226)   get_triples :232-297
 *   get_inv_mats :503-531   gel_energy :674-699   gel_forces :702-750
 *   duplicate_forces :754-779   affine_displace :783-793   dilate :796-806
 *   run_fire_md :865-1066   do_relaxation_run :1095-1391
 *   parse_batch_file :413-500
 */
#include "oracle.h"
#include "internal.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

void or_mesh_free(or_mesh *m)
{
    or_free(&m->net);
    free(m->facet);
    free(m->area);
    free(m->inv);
    free(m->pair);
    memset(m, 0, sizeof(*m));
}

static int or_pt_equiv(const double *pos, int a, double x, double y)
{
    return fabs(pos[2 * a] - x) <= OR_FLT_TOL && fabs(pos[2 * a + 1] - y) <= OR_FLT_TOL;
}

static const double *or_sort_pos;
static int or_cmp_xy(const void *pa, const void *pb)
{
    /* PointComp (:55-62): x first with a 1e-10 tolerance, then y */
    int a = *(const int *)pa, b = *(const int *)pb;
    const double *p = or_sort_pos;
    if (p[2 * a] < p[2 * b] - OR_FLT_TOL) return -1;
    if (p[2 * a] > p[2 * b] + OR_FLT_TOL) return 1;
    if (p[2 * a + 1] < p[2 * b + 1] - OR_FLT_TOL) return -1;
    if (p[2 * a + 1] > p[2 * b + 1] + OR_FLT_TOL) return 1;
    return a < b ? -1 : (a > b);
}

/* Twin pairs (:219-226): every point (the first of coincident ones) with
 * x < min_x + 1e-10, in (x, y) order, paired with the first point at
 * (x + offset, y) within 1e-10 in each coordinate. */
static int or_mesh_pairs(or_mesh *m)
{
    const double *pos = m->net.pos;
    int n = m->net.n, i, nl = 0, nr = 0, k, j;
    int *left = (int *)malloc(sizeof(int) * (size_t)n);
    int *right = (int *)malloc(sizeof(int) * (size_t)n);
    m->pair = (int *)malloc(sizeof(int) * 2 * (size_t)n);
    m->np = 0;
    if (!left || !right || !m->pair) { free(left); free(right); return OR_E_ALLOC; }
    for (i = 0; i < n; i++) {
        double x = pos[2 * i];
        if (x < m->minx + OR_FLT_TOL) left[nl++] = i;
        if (x >= m->minx + m->offset - 2 * OR_FLT_TOL && x <= m->minx + m->offset + 3 * OR_FLT_TOL)
            right[nr++] = i;
    }
    or_sort_pos = pos;
    qsort(left, (size_t)nl, sizeof(int), or_cmp_xy);
    for (k = 0; k < nl; k++) {
        int a = left[k], dup = 0, q = -1;
        double x = pos[2 * a] + m->offset, y = pos[2 * a + 1];
        for (j = 0; j < nl && !dup; j++) /* coincident with an earlier point */
            dup = left[j] < a && or_pt_equiv(pos, left[j], pos[2 * a], pos[2 * a + 1]);
        if (dup) continue;
        for (j = 0; j < nr; j++)
            if (or_pt_equiv(pos, right[j], x, y) && (q < 0 || right[j] < q)) q = right[j];
        if (q >= 0) {
            m->pair[2 * m->np] = a;
            m->pair[2 * m->np + 1] = q;
            m->np++;
        }
    }
    free(left);
    free(right);
    return OR_OK;
}

/* read_network_mesh :72-230.  Header "%d %d %d %lf" = points, edges,
 * facets, offset; then points, edges "%d %d %hhd", facets "%d %d %d" and
 * one positive area per facet, each section taking the lines that parse. */
int or_read_mesh(const char *path, or_mesh *m)
{
    FILE *fp;
    or_line ln = {0, 0};
    int eof = 0, n, e, nf, num_read, i1, i2, i3, nb = 0, np = 0, k = 0, na = 0;
    signed char wx;
    double W, x, y, a, minx = FLT_MAX, miny = FLT_MAX, maxx = FLT_MIN, maxy = FLT_MIN;
    int rc = OR_OK;
    or_net *net = &m->net;

    memset(m, 0, sizeof(*m));
    fp = fopen(path, "r");
    if (!fp) return OR_E_OPEN;
    if (or_getline(fp, &ln, &eof)) { fclose(fp); return OR_E_ALLOC; }
    num_read = sscanf(ln.buf, "%d %d %d %lf", &n, &e, &nf, &W);
    if (num_read < 4 || W < 0 || n <= 0 || e < 0 || nf < 0) {
        fclose(fp); free(ln.buf); return OR_E_HEADER;
    }
    net->n = n; net->width = W; net->height = 0.0;
    m->offset = W;
    m->nf = nf;
    net->pos = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    net->bond = (int *)malloc(sizeof(int) * 2 * (size_t)(e ? e : 1));
    net->rest = (double *)malloc(sizeof(double) * (size_t)(e ? e : 1));
    net->bwrap = (signed char *)malloc(2 * (size_t)(e ? e : 1));
    net->fcode = (unsigned char *)malloc((size_t)n);
    m->facet = (int *)malloc(sizeof(int) * 3 * (size_t)(nf ? nf : 1));
    m->area = (double *)malloc(sizeof(double) * (size_t)(nf ? nf : 1));
    m->inv = (double *)malloc(sizeof(double) * 6 * (size_t)(nf ? nf : 1));
    if (!net->pos || !net->bond || !net->rest || !net->bwrap || !net->fcode || !m->facet ||
        !m->area || !m->inv) {
        rc = OR_E_ALLOC; goto out;
    }
    while (!eof && np < n) {
        or_getline(fp, &ln, &eof);
        if (sscanf(ln.buf, "%lf %lf", &x, &y) == 2) {
            net->pos[2 * np] = x;
            net->pos[2 * np + 1] = y;
            if (x < minx) minx = x;
            if (y < miny) miny = y;
            if (x > maxx) maxx = x;
            if (y > maxy) maxy = y;
            np++;
        }
    }
    if (eof) { rc = OR_E_POINTS; goto out; }                    /* :114-119 */
    while (!eof && nb < e) {                                      /* :122-141 */
        or_getline(fp, &ln, &eof);
        if (sscanf(ln.buf, "%d %d %hhd", &i1, &i2, &wx) == 3) {
            if ((i1 < i2 ? i1 : i2) > -1 && (i1 > i2 ? i1 : i2) < n) {
                net->bond[2 * nb] = i1;
                net->bond[2 * nb + 1] = i2;
                net->bwrap[2 * nb] = wx;
                net->bwrap[2 * nb + 1] = 0;
                net->rest[nb] = or_rest_length(net->pos, W, 0.0, i1, i2, wx, 0);
                nb++;
            } else {
                rc = OR_E_INDEX; goto out;
            }
        }
    }
    net->e = nb;
    if (nb < e) { rc = OR_E_EDGES; goto out; }                    /* :145-150 */
    if (eof && nf > 0) { rc = OR_E_FACETS; goto out; }            /* :154-159 */
    while (!eof && k < nf) {                                      /* :162-180 */
        or_getline(fp, &ln, &eof);
        if (sscanf(ln.buf, "%d %d %d", &i1, &i2, &i3) == 3) {
            if (i1 < 0 || i1 >= n || i2 < 0 || i2 >= n || i3 < 0 || i3 >= n) {
                rc = OR_E_INDEX; goto out;
            }
            m->facet[3 * k] = i1;
            m->facet[3 * k + 1] = i2;
            m->facet[3 * k + 2] = i3;
            k++;
        }
    }
    if (k < nf) { rc = OR_E_FACETS; goto out; }                   /* :184-189 */
    if (eof && nf > 0) { rc = OR_E_FACETS; goto out; }            /* :193-198 */
    while (!eof && na < nf) {                                     /* :201-212 */
        or_getline(fp, &ln, &eof);
        if (sscanf(ln.buf, "%lf", &a) == 1) {
            if (!(a > 0)) { rc = OR_E_FACETS; goto out; }
            m->area[na++] = a;
        }
    }
    if (na < nf) rc = OR_E_FACETS;                                /* :216-219 */
out:
    fclose(fp);
    free(ln.buf);
    net->miny = miny;
    net->maxy = maxy;
    m->minx = minx;
    m->maxx = maxx;
    if (rc == OR_OK) {
        or_set_clamps(net);
        or_mesh_inv_mats(m, net->pos);
        rc = or_mesh_pairs(m);
    }
    if (rc != OR_OK) or_mesh_free(m);
    return rc;
}

/* get_inv_mats :503-531 on the given geometry: rows k = 0..2 of the inverse
 * of [x; y; 1] restricted to its first two columns, so that
 * F = [x1 x2 x3; y1 y2 y3] M is the deformation gradient. */
void or_mesh_inv_mats(or_mesh *m, const double *pos)
{
    int f;
    for (f = 0; f < m->nf; f++) {
        const int *t = m->facet + 3 * f;
        double x1 = pos[2 * t[0]], y1 = pos[2 * t[0] + 1];
        double x2 = pos[2 * t[1]], y2 = pos[2 * t[1] + 1];
        double x3 = pos[2 * t[2]], y3 = pos[2 * t[2] + 1];
        double det = fma(x1, y2 - y3, fma(x2, y3 - y1, x3 * (y1 - y2)));
        double *M = m->inv + 6 * f;
        M[0] = (y2 - y3) / det;
        M[1] = (x3 - x2) / det;
        M[2] = (y3 - y1) / det;
        M[3] = (x1 - x3) / det;
        M[4] = (y1 - y2) / det;
        M[5] = (x2 - x1) / det;
    }
}

/* Deformation gradient of facet f at pos. */
static void or_defgrad(const or_mesh *m, int f, const double *pos, double F[4])
{
    const int *t = m->facet + 3 * f;
    const double *M = m->inv + 6 * f;
    double x1 = pos[2 * t[0]], y1 = pos[2 * t[0] + 1];
    double x2 = pos[2 * t[1]], y2 = pos[2 * t[1] + 1];
    double x3 = pos[2 * t[2]], y3 = pos[2 * t[2] + 1];
    F[0] = fma(x3, M[4], fma(x2, M[2], x1 * M[0]));
    F[1] = fma(x3, M[5], fma(x2, M[3], x1 * M[1]));
    F[2] = fma(y3, M[4], fma(y2, M[2], y1 * M[0]));
    F[3] = fma(y3, M[5], fma(y2, M[3], y1 * M[1]));
}

/* gel_energy :674-699: A [smod (|F|^2 / J - 2) + bmod (J - 1 - d)^2] / 2 */
double or_gel_energy(const or_mesh *m, double smod, double bmod, double pd, const double *pos)
{
    int f;
    double sum = 0.0;
    for (f = 0; f < m->nf; f++) {
        double F[4], J, I1, A = m->area[f], u;
        or_defgrad(m, f, pos, F);
        J = fma(F[0], F[3], -(F[1] * F[2]));
        I1 = fma(F[0], F[0], fma(F[1], F[1], fma(F[2], F[2], F[3] * F[3])));
        u = smod * A * (I1 / J - 2.0) / 2.0;
        u += bmod * A * (J - 1.0 - pd) * (J - 1.0 - pd) / 2.0;
        sum += u;
    }
    return sum;
}

/* gel_forces :702-750: -dU/dx of each facet, masked per DOF by the clamp
 * code of the receiving point; then duplicate_forces :754-779: each twin
 * receives the sum of the two (unmasked, as the reference adds them). */
static void or_gel_forces(const or_mesh *m, double smod, double bmod, double pd,
                          const double *pos, double *force)
{
    int f, k;
    const unsigned char *fc = m->net.fcode;
    for (f = 0; f < m->nf; f++) {
        const int *t = m->facet + 3 * f;
        const double *M = m->inv + 6 * f;
        double F[4], J, I1, A = m->area[f], c1, c2, d11, d12, d21, d22;
        or_defgrad(m, f, pos, F);
        J = fma(F[0], F[3], -(F[1] * F[2]));
        I1 = fma(F[0], F[0], fma(F[1], F[1], fma(F[2], F[2], F[3] * F[3])));
        c1 = -smod * A / J;
        c2 = (bmod * (pd + 1.0 - J) + smod * I1 / (2.0 * J * J)) * A;
        d11 = fma(F[0], c1, F[3] * c2);
        d12 = fma(F[1], c1, -(F[2] * c2));
        d21 = fma(F[2], c1, -(F[1] * c2));
        d22 = fma(F[3], c1, F[0] * c2);
        for (k = 0; k < 3; k++) {
            int v = t[k];
            double fx = fma(d12, M[2 * k + 1], d11 * M[2 * k]);
            double fy = fma(d22, M[2 * k + 1], d21 * M[2 * k]);
            if (fc[v] & 1) force[2 * v] += fx;
            if (fc[v] & 2) force[2 * v + 1] += fy;
        }
    }
}

static void or_twin_sum(const or_mesh *m, double *force)
{
    int p;
    for (p = 0; p < m->np; p++) {
        int a = m->pair[2 * p], b = m->pair[2 * p + 1];
        double sx = force[2 * a] + force[2 * b], sy = force[2 * a + 1] + force[2 * b + 1];
        force[2 * a] = sx;
        force[2 * b] = sx;
        force[2 * a + 1] = sy;
        force[2 * b + 1] = sy;
    }
}

/* Force of run_fire_md :936-951 / :969-985: bonds (ks > 0), triples
 * (kb > 0), then, only when bmod > 0, the gel and the twin sum. */
void or_mesh_forces(const or_mesh *m, double ks, double kb, double smod, double bmod, double pd,
                    const double *pos, double *force)
{
    or_forces(&m->net, ks, kb, pos, force);
    if (bmod > 0) {
        or_gel_forces(m, smod, bmod, pd, pos, force);
        or_twin_sum(m, force);
    }
}

/* Energies of run_fire_md :1030-1056: {E_stretch, E_bend, E_gel}, each 0
 * when its modulus is 0 (the gel when bmod is 0). */
void or_mesh_energies(const or_mesh *m, double ks, double kb, double smod, double bmod,
                      double pd, const double *pos, double e[3])
{
    or_energies(&m->net, ks, kb, pos, &e[0], &e[1]);
    if (!(ks > 0)) e[0] = 0.0;
    if (!(kb > 0)) e[1] = 0.0;
    e[2] = bmod > 0 ? or_gel_energy(m, smod, bmod, pd, pos) : 0.0;
}

typedef struct {
    const or_mesh *m;
    double ks, kb, smod, bmod, pd;
} or_mesh_field;

static void or_mesh_force_cb(const void *ctx, const double *pos, double *force)
{
    const or_mesh_field *c = (const or_mesh_field *)ctx;
    or_mesh_forces(c->m, c->ks, c->kb, c->smod, c->bmod, c->pd, pos, force);
}

/* run_fire_md :865-1066 with the composite force; report: [0] rms |F|,
 * [1] E_stretch, [2] E_bend, [3] E_gel, [4] max node |F|. */
int or_mesh_fire(const or_mesh *m, double ks, double kb, double smod, double bmod, double pd,
                 double *pos, const double fp[5], int nmin, double mass, double fcut,
                 int max_steps, double report[5], int *steps)
{
    or_mesh_field c = {m, ks, kb, smod, bmod, pd};
    int rc = or_fire_core(m->net.n, or_mesh_force_cb, &c, pos, fp, nmin, mass, fcut, max_steps,
                          OR_CRIT_RMS, &report[0], &report[4], steps);
    if (rc == OR_E_ALLOC) return rc;
    or_mesh_energies(m, ks, kb, smod, bmod, pd, pos, &report[1]);
    return rc;
}

/* Set-up of do_relaxation_run :1138-1275 after reading: the box offset
 * grows by s = sqrt(1 + dilation) (:1141) and the triples are found with
 * that offset on the undilated points (:1143-1145); clamp codes and inverse
 * matrices come from the undilated geometry (:1221-1250); the points are
 * dilated by s about (min_x, min_y) (:1254); the network's stretching energy
 * in that state gives its bulk modulus 2 E / (area dilation^2) (:1256-1271)
 * and the gel's preferred dilation d (bmod + net_bmod) / bmod (:1274).
 * pos (2n) receives the dilated points.  With dilation 0 the preferred
 * dilation is 0 (the reference divides 0 by 0 there). */
int or_mesh_setup(or_mesh *m, double ks, double kb, double bmod, double dilation, double *pos,
                  double *pref_d)
{
    int i, rc;
    double s = sqrt(1.0 + dilation), area, es, eb, net_bmod = 0.0;
    m->net.width = m->offset * s;
    if (kb > 0) {
        rc = or_get_triples(&m->net);
        if (rc) return rc;
    }
    or_mesh_inv_mats(m, m->net.pos);
    for (i = 0; i < m->net.n; i++) {
        pos[2 * i] = fma(s, m->net.pos[2 * i] - m->minx, m->minx);
        pos[2 * i + 1] = fma(s, m->net.pos[2 * i + 1] - m->net.miny, m->net.miny);
    }
    area = (m->maxx - m->minx) * (m->net.maxy - m->net.miny);
    if (ks > 0 && dilation > 0) {
        or_energies(&m->net, ks, 0.0, pos, &es, &eb);
        net_bmod = es * (2.0 / area / (dilation * dilation));
    }
    *pref_d = bmod > 0 ? dilation * (bmod + net_bmod) / bmod : dilation;
    return OR_OK;
}

/* parse_batch_file :413-500: line 2 is "ks kb bmod [smod [dilation]]"
 * (smod and dilation 0 when absent); the rest as the network program's. */
int or_parse_mesh_batch(const char *path, char *base, int base_cap, char *tag, int tag_cap,
                        or_mesh_batch *b)
{
    or_batch nb;
    FILE *fp;
    or_line ln = {0, 0};
    int eof = 0, got = 0, rc;
    memset(b, 0, sizeof(*b));
    rc = or_parse_batch(path, base, base_cap, tag, tag_cap, &nb);
    if (rc) return rc;
    fp = fopen(path, "r");
    if (!fp) return OR_E_OPEN;
    while (!eof && got < 2) { /* the second non-blank, non-comment line */
        or_getline(fp, &ln, &eof);
        if (ln.buf[0] == 0 || ln.buf[0] == '#') continue;
        got++;
    }
    rc = sscanf(ln.buf, "%lf %lf %lf %lf %lf", &b->ks, &b->kb, &b->bmod, &b->smod, &b->dilation);
    fclose(fp);
    free(ln.buf);
    if (rc < 3 || b->ks < 0 || b->kb < 0 || b->bmod < 0 || b->smod < 0 || b->dilation < 0)
        return OR_E_ARG;
    memcpy(b->fp, nb.fp, sizeof(b->fp));
    b->nmin = nb.nmin; b->mass = nb.mass; b->fcut = nb.fcut; b->nmax = nb.nmax;
    b->cmin = nb.cmin; b->cmax = nb.cmax; b->cinc = nb.cinc;
    b->smin = nb.smin; b->smax = nb.smax; b->sinc = nb.sinc;
    return OR_OK;
}

static void or_mesh_log(FILE *lf, const double rep[5], int steps, double fcut)
{
    if (!lf) return;
    fprintf(lf, rep[0] <= fcut ? "Convergence reached: " : "Convergence not reached: ");
    fprintf(lf, "Steps: %d, RMS Force: %12.8le\n", steps, rep[0]);
}

/* do_relaxation_run :1095-1391 for one composite: set-up, then the
 * compression ramp (the first step applies the minimum, each later one the
 * increment, the box height shrinking by each), then the shear ramp, a
 * relax after each step; log, edata (6 columns) and .pos outputs.
 * pref_d_out (optional) receives the preferred dilation. */
int or_run_mesh_protocol(const char *net_file, const char *log_name, const char *edata_name,
                         const char *disp_base, const or_mesh_batch *b, int strain_mode,
                         double *pref_d_out)
{
    or_mesh m;
    double *pos, range, smult, mn, mx, inc, cur = 0.0, max_comp, rep[5], pd;
    int rc, steps, num_digits, disp_count = 1;
    FILE *lf, *ef;
    clock_t ct = clock();
    rc = or_read_mesh(net_file, &m);
    if (rc) return rc;
    if ((b->cmax > 0 && b->cinc <= 0) || (b->smax > 0 && b->sinc <= 0)) {
        or_mesh_free(&m);
        return OR_E_ARG;
    }
    pos = (double *)malloc(sizeof(double) * 2 * (size_t)m.net.n);
    if (!pos) { or_mesh_free(&m); return OR_E_ALLOC; }
    rc = or_mesh_setup(&m, b->ks, b->kb, b->bmod, b->dilation, pos, &pd);
    if (rc) { free(pos); or_mesh_free(&m); return rc; }
    if (pref_d_out) *pref_d_out = pd;
    range = (m.net.maxy - m.net.miny) * sqrt(1.0 + b->dilation);
    smult = strain_mode ? range : 1.0;
    num_digits = or_dec_digits(1 + (b->sinc > 0 ? (int)((b->smax - b->smin) / b->sinc) : 0));
    lf = fopen(log_name, "w");
    ef = fopen(edata_name, "w");
    mn = b->cmin * smult; mx = b->cmax * smult; inc = b->cinc * smult;
    if (mx > 0) {
        or_affine(&m.net, pos, m.net.miny, range, 0.0, -mn);
        range -= mn;
        for (cur = mn; cur <= mx + OR_FLT_TOL; cur += inc) {
            or_mesh_fire(&m, b->ks, b->kb, b->smod, b->bmod, pd, pos, b->fp, b->nmin, b->mass,
                         b->fcut, b->nmax, rep, &steps);
            or_mesh_log(lf, rep, steps, b->fcut);
            if (ef) fprintf(ef, "%12.0lf\t%12.10lf\t%12.8le\t%12.8le\t%12.8le\t%12.8le\n", 0.,
                            cur, rep[1], rep[2], rep[3], rep[1] + rep[2] + rep[3]);
            if (cur < mx) {
                or_affine(&m.net, pos, m.net.miny, range, 0.0, -inc);
                range -= inc;
            }
        }
    }
    max_comp = mx > mn ? cur - inc : mx;
    mn = b->smin * smult; mx = b->smax * smult; inc = b->sinc * smult;
    if (mx > 0) {
        or_affine(&m.net, pos, m.net.miny, range, mn, 0.0);
        for (cur = mn; cur <= mx + OR_FLT_TOL; cur += inc) {
            char name[4096];
            FILE *pf;
            int k;
            or_mesh_fire(&m, b->ks, b->kb, b->smod, b->bmod, pd, pos, b->fp, b->nmin, b->mass,
                         b->fcut, b->nmax, rep, &steps);
            or_mesh_log(lf, rep, steps, b->fcut);
            if (ef) fprintf(ef, "%12.8lf\t%12.8lf\t%12.8le\t%12.8le\t%12.8le\t%12.8le\n", cur,
                            max_comp, rep[1], rep[2], rep[3], rep[1] + rep[2] + rep[3]);
            or_report_name(disp_base, num_digits, disp_count++, "pos", name, sizeof(name));
            pf = fopen(name, "w");
            if (pf) {
                for (k = 0; k < m.net.n; k++)
                    fprintf(pf, "%2.10le\t%2.10le\n", pos[2 * k], pos[2 * k + 1]);
                fclose(pf);
            }
            if (cur < mx) or_affine(&m.net, pos, m.net.miny, range, inc, 0.0);
        }
    }
    ct = clock() - ct;
    if (lf) {
        fprintf(lf, "Elapsed time: %f seconds\n", ((float)ct) / CLOCKS_PER_SEC);
        fclose(lf);
    }
    if (ef) fclose(ef);
    free(pos);
    or_mesh_free(&m);
    return OR_OK;
}
