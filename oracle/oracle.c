/*
 * CPU FP64 oracle for the network relaxation step.  TEST INFRASTRUCTURE ONLY:
 * used by tests/, the smoke check and bench.py's CPU-baseline leg as the
 * checker, never by the product.  See oracle.h for the map onto the
 * reference's functions.
 */
#include "oracle.h"
#include "internal.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------------ */
/* std::getline emulation: returns the line without '\n'; *eof is set when   */
/* end of file was reached while reading (no terminating newline).           */
/* ------------------------------------------------------------------------ */
int or_getline(FILE *fp, or_line *ln, int *eof)
{
    size_t len = 0;
    int c;
    *eof = 0;
    if (!ln->buf) {
        ln->cap = 256;
        ln->buf = (char *)malloc(ln->cap);
        if (!ln->buf) return -1;
    }
    for (;;) {
        c = fgetc(fp);
        if (c == EOF) { *eof = 1; break; }
        if (c == '\n') break;
        if (len + 2 > ln->cap) {
            char *nb = (char *)realloc(ln->buf, ln->cap * 2);
            if (!nb) return -1;
            ln->buf = nb;
            ln->cap *= 2;
        }
        ln->buf[len++] = (char)c;
    }
    ln->buf[len] = 0;
    return 0;
}

void or_free(or_net *net)
{
    free(net->pos); free(net->bond); free(net->rest); free(net->bwrap);
    free(net->tri); free(net->twrap); free(net->tkb); free(net->fcode);
    free(net->tbond);
    memset(net, 0, sizeof(*net));
}

/* Rest length of bond (i, j, wx, wy) in the input geometry:
 * large_strain_relax.cu:97-99, dx = x_j + W*w - x_i, |r| = sqrt(dx*dx+dy*dy).
 * The y image term is the Lees-Edwards extension; it is +0.0 in reference mode,
 * which leaves the reference's value unchanged. */
double or_rest_length(const double *pos, double W, double H, int i,
                             int j, int wx, int wy)
{
    double dx = pos[2 * j] + W * wx - pos[2 * i];
    double dy = pos[2 * j + 1] + H * wy - pos[2 * i + 1];
    double a = dx * dx;
    double b = dy * dy;
    return sqrt(a + b);
}

/* read_network, large_strain_relax.cu:42-114.  Extensions: a 4th header
 * number is the box height (Lees-Edwards mode) and, in that mode only, a 4th
 * integer on an edge line is the y wrap.  Unlike the reference's caller
 * (:829), a short edge list is an error. */
int or_read_network(const char *path, or_net *net)
{
    FILE *fp;
    or_line ln = {0, 0};
    int eof = 0, n, e, num_read, i1, i2, nb = 0, np = 0;
    signed char wx, wy;
    double W, H = 0.0, x, y, miny = FLT_MAX, maxy = FLT_MIN;
    int rc = OR_OK;

    memset(net, 0, sizeof(*net));
    fp = fopen(path, "r");
    if (!fp) return OR_E_OPEN;
    if (or_getline(fp, &ln, &eof)) { fclose(fp); return OR_E_ALLOC; }
    num_read = sscanf(ln.buf, "%d %d %lf %lf", &n, &e, &W, &H);
    if (num_read < 3 || W < 0 || n <= 0 || e < 0) {
        fclose(fp); free(ln.buf); return OR_E_HEADER;
    }
    if (num_read < 4 || H < 0) H = 0.0;
    net->n = n; net->width = W; net->height = H;
    net->pos = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    net->bond = (int *)malloc(sizeof(int) * 2 * (size_t)(e ? e : 1));
    net->rest = (double *)malloc(sizeof(double) * (size_t)(e ? e : 1));
    net->bwrap = (signed char *)malloc(2 * (size_t)(e ? e : 1));
    net->fcode = (unsigned char *)malloc((size_t)n);
    if (!net->pos || !net->bond || !net->rest || !net->bwrap || !net->fcode) {
        rc = OR_E_ALLOC; goto out;
    }
    /* points: lines that do not hold two numbers are skipped (:68-77) */
    while (!eof && np < n) {
        or_getline(fp, &ln, &eof);
        if (sscanf(ln.buf, "%lf %lf", &x, &y) == 2) {
            net->pos[2 * np] = x;
            net->pos[2 * np + 1] = y;
            if (y < miny) miny = y;
            if (y > maxy) maxy = y;
            np++;
        }
    }
    if (eof) { rc = OR_E_POINTS; goto out; }          /* :81-85 */
    /* edges (:88-107) */
    while (!eof && nb < e) {
        int wxi, wyi = 0;
        or_getline(fp, &ln, &eof);
        num_read = sscanf(ln.buf, "%d %d %hhd %hhd", &i1, &i2, &wx, &wy);
        if (num_read >= 3) {
            if ((i1 < i2 ? i1 : i2) > -1 && (i1 > i2 ? i1 : i2) < n) {
                wxi = wx;
                if (H > 0 && num_read == 4) wyi = wy;
                net->bond[2 * nb] = i1;
                net->bond[2 * nb + 1] = i2;
                net->bwrap[2 * nb] = (signed char)wxi;
                net->bwrap[2 * nb + 1] = (signed char)wyi;
                net->rest[nb] = or_rest_length(net->pos, W, H, i1, i2, wxi, wyi);
                nb++;
            } else {
                rc = OR_E_INDEX; goto out;
            }
        }
    }
    net->e = nb;
    if (nb < e) rc = OR_E_EDGES;
out:
    fclose(fp);
    free(ln.buf);
    net->miny = miny;
    net->maxy = maxy;
    if (rc == OR_OK) or_set_clamps(net);
    else or_free(net);
    return rc;
}

int or_net_from_arrays(int n, int e, double width, double height,
                       const double *pos, const int *bond,
                       const signed char *bwrap, or_net *net)
{
    int i;
    double miny = FLT_MAX, maxy = FLT_MIN;
    memset(net, 0, sizeof(*net));
    if (n <= 0 || e < 0 || width < 0 || height < 0) return OR_E_ARG;
    net->n = n; net->e = e; net->width = width; net->height = height;
    net->pos = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    net->bond = (int *)malloc(sizeof(int) * 2 * (size_t)(e ? e : 1));
    net->rest = (double *)malloc(sizeof(double) * (size_t)(e ? e : 1));
    net->bwrap = (signed char *)malloc(2 * (size_t)(e ? e : 1));
    net->fcode = (unsigned char *)malloc((size_t)n);
    if (!net->pos || !net->bond || !net->rest || !net->bwrap || !net->fcode) {
        or_free(net);
        return OR_E_ALLOC;
    }
    memcpy(net->pos, pos, sizeof(double) * 2 * (size_t)n);
    for (i = 0; i < n; i++) {
        double y = pos[2 * i + 1];
        if (y < miny) miny = y;
        if (y > maxy) maxy = y;
    }
    for (i = 0; i < e; i++) {
        int a = bond[2 * i], b = bond[2 * i + 1];
        int wx = bwrap[2 * i], wy = height > 0 ? bwrap[2 * i + 1] : 0;
        if (a < 0 || b < 0 || a >= n || b >= n) { or_free(net); return OR_E_INDEX; }
        net->bond[2 * i] = a; net->bond[2 * i + 1] = b;
        net->bwrap[2 * i] = (signed char)wx; net->bwrap[2 * i + 1] = (signed char)wy;
        net->rest[i] = or_rest_length(net->pos, width, height, a, b, wx, wy);
    }
    net->miny = miny; net->maxy = maxy;
    or_set_clamps(net);
    return OR_OK;
}

/* Clamp codes, large_strain_relax.cu:896-902: a node within 1e-10 of the
 * lowest or highest input y has both DOFs frozen.  In Lees-Edwards mode the
 * box is periodic in y and nothing is clamped. */
void or_set_clamps(or_net *net)
{
    int i;
    for (i = 0; i < net->n; i++) {
        double y = net->pos[2 * i + 1];
        if (net->height > 0) net->fcode[i] = 3;
        else if (y < net->miny + OR_FLT_TOL || y > net->maxy - OR_FLT_TOL)
            net->fcode[i] = 0;
        else net->fcode[i] = 3;
    }
}

/* ------------------------------------------------------------------------ */
/* get_triples, large_strain_relax.cu:116-184.                               */
/* Adjacency lists per node in edge-file order, the reverse direction with    */
/* negated wrap (:126-139); hinges in ascending node order (std::map order,   */
/* :147); neighbour pairs i<j in list order; antiparallel iff dot < 0 and     */
/* |dot^2 - |a|^2|b|^2| < 1e-10 in the input geometry (:176).                 */
/* Also records the two bonds of each triple (for fiber labels).             */
/* ------------------------------------------------------------------------ */

static int or_build_adjacency(const or_net *net, int **off_out, int **ent_out)
{
    int *deg = (int *)calloc((size_t)net->n + 1, sizeof(int));
    int *ent, *fill, i;
    if (!deg) return OR_E_ALLOC;
    for (i = 0; i < net->e; i++) {
        deg[net->bond[2 * i] + 1]++;
        deg[net->bond[2 * i + 1] + 1]++;
    }
    for (i = 0; i < net->n; i++) deg[i + 1] += deg[i];
    ent = (int *)malloc(sizeof(int) * 2 * (size_t)(net->e ? net->e : 1) * 2);
    fill = (int *)malloc(sizeof(int) * (size_t)net->n);
    if (!ent || !fill) { free(deg); free(ent); free(fill); return OR_E_ALLOC; }
    memcpy(fill, deg, sizeof(int) * (size_t)net->n);
    /* entry = (bond index, direction); direction 0: i->j, 1: j->i */
    for (i = 0; i < net->e; i++) {
        int a = net->bond[2 * i], b = net->bond[2 * i + 1];
        int k = fill[a]++;
        ent[2 * k] = i; ent[2 * k + 1] = 0;
        k = fill[b]++;
        ent[2 * k] = i; ent[2 * k + 1] = 1;
    }
    free(fill);
    *off_out = deg;
    *ent_out = ent;
    return OR_OK;
}

int or_get_triples(or_net *net)
{
    int *off = NULL, *ent = NULL, h, cap = 64, t = 0, rc;
    double W = net->width, H = net->height;
    int *tb;
    free(net->tri); free(net->twrap); free(net->tkb); free(net->tbond);
    net->tri = NULL; net->twrap = NULL; net->tkb = NULL; net->tbond = NULL;
    rc = or_build_adjacency(net, &off, &ent);
    if (rc) return rc;
    net->tri = (int *)malloc(sizeof(int) * 3 * (size_t)cap);
    net->twrap = (signed char *)malloc(4 * (size_t)cap);
    tb = (int *)malloc(sizeof(int) * 2 * (size_t)cap);
    net->tbond = tb;
    if (!net->tri || !net->twrap || !tb) { rc = OR_E_ALLOC; goto out; }
    for (h = 0; h < net->n; h++) {
        int i, j;
        double x = net->pos[2 * h], y = net->pos[2 * h + 1];
        for (i = off[h]; i < off[h + 1]; i++) {
            int b2 = ent[2 * i], d2 = ent[2 * i + 1];
            int n2 = net->bond[2 * b2 + (d2 ? 0 : 1)];
            int wx2 = d2 ? -net->bwrap[2 * b2] : net->bwrap[2 * b2];
            int wy2 = d2 ? -net->bwrap[2 * b2 + 1] : net->bwrap[2 * b2 + 1];
            double dx2 = net->pos[2 * n2] + W * wx2 - x;
            double dy2 = net->pos[2 * n2 + 1] + H * wy2 - y;
            for (j = i + 1; j < off[h + 1]; j++) {
                int b3 = ent[2 * j], d3 = ent[2 * j + 1];
                int n3 = net->bond[2 * b3 + (d3 ? 0 : 1)];
                int wx3 = d3 ? -net->bwrap[2 * b3] : net->bwrap[2 * b3];
                int wy3 = d3 ? -net->bwrap[2 * b3 + 1] : net->bwrap[2 * b3 + 1];
                double dx3 = net->pos[2 * n3] + W * wx3 - x;
                double dy3 = net->pos[2 * n3 + 1] + H * wy3 - y;
                double p1 = dx2 * dx3, p2 = dy2 * dy3;
                double dot = p1 + p2;
                double m1 = dx2 * dx2, m2 = dy2 * dy2, m3 = dx3 * dx3, m4 = dy3 * dy3;
                double mag_sq = (m1 + m2) * (m3 + m4);
                double dd = dot * dot;
                if (dot < 0 && fabs(dd - mag_sq) < OR_FLT_TOL) {
                    if (t == cap) {
                        cap *= 2;
                        net->tri = (int *)realloc(net->tri, sizeof(int) * 3 * (size_t)cap);
                        net->twrap = (signed char *)realloc(net->twrap, 4 * (size_t)cap);
                        tb = (int *)realloc(tb, sizeof(int) * 2 * (size_t)cap);
                        net->tbond = tb;
                        if (!net->tri || !net->twrap || !tb) { rc = OR_E_ALLOC; goto out; }
                    }
                    net->tri[3 * t] = h; net->tri[3 * t + 1] = n2; net->tri[3 * t + 2] = n3;
                    net->twrap[4 * t] = (signed char)wx2;
                    net->twrap[4 * t + 1] = (signed char)wy2;
                    net->twrap[4 * t + 2] = (signed char)wx3;
                    net->twrap[4 * t + 3] = (signed char)wy3;
                    tb[2 * t] = b2;
                    tb[2 * t + 1] = b3;
                    t++;
                }
            }
        }
    }
    net->t = t;
out:
    free(off);
    free(ent);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Fibers and per-fiber bending stiffness (extension; no reference code).    */
/* A fiber is a connected set of bonds joined by triples; its label is the    */
/* smallest bond index in it.  kappa = exp(ln klo + u*(ln khi - ln klo)) with */
/* u = or_uniform(seed, label).                                               */
/* ------------------------------------------------------------------------ */
static int or_find(int *p, int a)
{
    while (p[a] != a) { p[a] = p[p[a]]; a = p[a]; }
    return a;
}

int or_fiber_labels(const or_net *net, int *label)
{
    int *par, i;
    if (!net->tbond && net->t > 0) return OR_E_ARG;
    par = (int *)malloc(sizeof(int) * (size_t)(net->e ? net->e : 1));
    if (!par) return OR_E_ALLOC;
    for (i = 0; i < net->e; i++) par[i] = i;
    for (i = 0; i < net->t; i++) {
        int a = or_find(par, net->tbond[2 * i]);
        int b = or_find(par, net->tbond[2 * i + 1]);
        if (a < b) par[b] = a; else if (b < a) par[a] = b;
    }
    for (i = 0; i < net->t; i++) label[i] = or_find(par, net->tbond[2 * i]);
    free(par);
    return OR_OK;
}

int or_set_fiber_kappa(or_net *net, double klo, double khi, uint64_t seed)
{
    int *lab, i, rc;
    double llo = log(klo), lhi = log(khi);
    free(net->tkb);
    net->tkb = NULL;
    if (net->t == 0) return OR_OK;
    lab = (int *)malloc(sizeof(int) * (size_t)net->t);
    net->tkb = (double *)malloc(sizeof(double) * (size_t)net->t);
    if (!lab || !net->tkb) { free(lab); return OR_E_ALLOC; }
    rc = or_fiber_labels(net, lab);
    if (rc) { free(lab); return rc; }
    for (i = 0; i < net->t; i++) {
        double u = or_uniform(seed, (uint64_t)lab[i]);
        net->tkb[i] = exp(llo + u * (lhi - llo));
    }
    free(lab);
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Forces and energies.                                                      */
/* ------------------------------------------------------------------------ */
static inline void or_image(const or_net *net, int wx, int wy, double *ox,
                            double *oy)
{
    *ox = net->width * wx + net->le_shift * wy;
    *oy = net->height * wy;
}

/* axial_forces :386-414 and bending_forces :417-469, executed sequentially
 * in element order (bonds, then triples), contributions masked per DOF. */
void or_forces(const or_net *net, double ks, double kb, const double *pos,
               double *force)
{
    int b, t, i;
    memset(force, 0, sizeof(double) * 2 * (size_t)net->n);
    if (ks > 0) {
        for (b = 0; b < net->e; b++) {
            int p = net->bond[2 * b], q = net->bond[2 * b + 1];
            double ox, oy, rx, ry, d2, inv, s, fx, fy;
            or_image(net, net->bwrap[2 * b], net->bwrap[2 * b + 1], &ox, &oy);
            rx = pos[2 * q] + ox - pos[2 * p];
            ry = pos[2 * q + 1] + oy - pos[2 * p + 1];
            d2 = fma(rx, rx, ry * ry);
            inv = 1.0 / sqrt(d2);
            s = ks * (1.0 - net->rest[b] * inv);
            fx = s * rx;
            fy = s * ry;
            force[2 * p] += fx;
            force[2 * p + 1] += fy;
            force[2 * q] += -fx;
            force[2 * q + 1] += -fy;
        }
    }
    if (kb > 0) {
        for (t = 0; t < net->t; t++) {
            int h = net->tri[3 * t], a = net->tri[3 * t + 1], c = net->tri[3 * t + 2];
            double oxa, oya, oxb, oyb, ax, ay, bx, by, ia, ib, cr, dt, dth, k;
            double f2x, f2y, f3x, f3y, g;
            or_image(net, net->twrap[4 * t], net->twrap[4 * t + 1], &oxa, &oya);
            or_image(net, net->twrap[4 * t + 2], net->twrap[4 * t + 3], &oxb, &oyb);
            ax = pos[2 * a] + oxa - pos[2 * h];
            ay = pos[2 * a + 1] + oya - pos[2 * h + 1];
            bx = pos[2 * c] + oxb - pos[2 * h];
            by = pos[2 * c + 1] + oyb - pos[2 * h + 1];
            ia = 1.0 / fma(ax, ax, ay * ay);
            ib = 1.0 / fma(bx, bx, by * by);
            cr = fma(ax, by, -(ay * bx));
            dt = fma(ax, bx, ay * by);
            dth = atan2(-cr, -dt);
            k = net->tkb ? net->tkb[t] : kb;
            g = k * dth;
            f2x = g * ia;
            f2y = f2x * ax;
            f2x = f2x * -ay;
            f3x = g * ib;
            f3y = -f3x * bx;
            f3x = f3x * by;
            force[2 * h] += -f2x - f3x;
            force[2 * h + 1] += -f2y - f3y;
            force[2 * a] += f2x;
            force[2 * a + 1] += f2y;
            force[2 * c] += f3x;
            force[2 * c + 1] += f3y;
        }
    }
    for (i = 0; i < net->n; i++) {
        unsigned char m = net->fcode[i];
        if (!(m & 1)) force[2 * i] = 0.0;
        if (!(m & 2)) force[2 * i + 1] = 0.0;
    }
}

/* axial_energies :472-490 + Dasum :743, bending_energies :493-523 + Dasum
 * :751 (the reference's off-by-one read at :502 is not reproduced). */
void or_energies(const or_net *net, double ks, double kb, const double *pos,
                 double *es, double *eb)
{
    int b, t;
    double s = 0.0, u = 0.0;
    if (ks > 0) {
        for (b = 0; b < net->e; b++) {
            int p = net->bond[2 * b], q = net->bond[2 * b + 1];
            double ox, oy, rx, ry, d, dl;
            or_image(net, net->bwrap[2 * b], net->bwrap[2 * b + 1], &ox, &oy);
            rx = pos[2 * q] + ox - pos[2 * p];
            ry = pos[2 * q + 1] + oy - pos[2 * p + 1];
            d = sqrt(fma(rx, rx, ry * ry));
            dl = d - net->rest[b];
            s += 0.5 * ks * dl * dl;
        }
    }
    if (kb > 0) {
        for (t = 0; t < net->t; t++) {
            int h = net->tri[3 * t], a = net->tri[3 * t + 1], c = net->tri[3 * t + 2];
            double oxa, oya, oxb, oyb, ax, ay, bx, by, cr, dt, dth, k;
            or_image(net, net->twrap[4 * t], net->twrap[4 * t + 1], &oxa, &oya);
            or_image(net, net->twrap[4 * t + 2], net->twrap[4 * t + 3], &oxb, &oyb);
            ax = pos[2 * a] + oxa - pos[2 * h];
            ay = pos[2 * a + 1] + oya - pos[2 * h + 1];
            bx = pos[2 * c] + oxb - pos[2 * h];
            by = pos[2 * c + 1] + oyb - pos[2 * h + 1];
            cr = fma(ax, by, -(ay * bx));
            dt = fma(ax, bx, ay * by);
            dth = atan2(-cr, -dt);
            k = net->tkb ? net->tkb[t] : kb;
            u += 0.5 * k * dth * dth;
        }
    }
    *es = s;
    *eb = u;
}

/* Shear virial W = dE/dgamma under the affine shear x += gamma (y - y0) of
 * every element's geometry (extension: the reference reports no stress).
 * Bond: ks (1 - L/|r|) r_x r_y.  Triple with arms a = r_a - r_h, b = r_c - r_h
 * and end forces f_a, f_c as in or_forces: -(f_a,x a_y + f_c,x b_y).  The
 * shear stress is W over the box area. */
double or_shear_virial(const or_net *net, double ks, double kb, const double *pos)
{
    int b, t;
    double w = 0.0;
    if (ks > 0) {
        for (b = 0; b < net->e; b++) {
            int p = net->bond[2 * b], q = net->bond[2 * b + 1];
            double ox, oy, rx, ry, d2, inv, s;
            or_image(net, net->bwrap[2 * b], net->bwrap[2 * b + 1], &ox, &oy);
            rx = pos[2 * q] + ox - pos[2 * p];
            ry = pos[2 * q + 1] + oy - pos[2 * p + 1];
            d2 = fma(rx, rx, ry * ry);
            inv = 1.0 / sqrt(d2);
            s = ks * (1.0 - net->rest[b] * inv);
            w += s * rx * ry;
        }
    }
    if (kb > 0) {
        for (t = 0; t < net->t; t++) {
            int h = net->tri[3 * t], a = net->tri[3 * t + 1], c = net->tri[3 * t + 2];
            double oxa, oya, oxb, oyb, ax, ay, bx, by, cr, dt, dth, k, g, f2x, f3x;
            or_image(net, net->twrap[4 * t], net->twrap[4 * t + 1], &oxa, &oya);
            or_image(net, net->twrap[4 * t + 2], net->twrap[4 * t + 3], &oxb, &oyb);
            ax = pos[2 * a] + oxa - pos[2 * h];
            ay = pos[2 * a + 1] + oya - pos[2 * h + 1];
            bx = pos[2 * c] + oxb - pos[2 * h];
            by = pos[2 * c + 1] + oyb - pos[2 * h + 1];
            cr = fma(ax, by, -(ay * bx));
            dt = fma(ax, bx, ay * by);
            dth = atan2(-cr, -dt);
            k = net->tkb ? net->tkb[t] : kb;
            g = k * dth;
            f2x = g * (1.0 / fma(ax, ax, ay * ay)) * -ay;
            f3x = g * (1.0 / fma(bx, bx, by * by)) * by;
            w -= f2x * ay + f3x * by;
        }
    }
    return w;
}

/* affine_displace :527-537 over the N nodes (the reference launches 2N
 * threads with a `>` guard and so writes past the array; only the in-bounds
 * part is reproduced). */
void or_affine(const or_net *net, double *pos, double miny, double yrange,
               double xdisp, double ydisp)
{
    int i;
    for (i = 0; i < net->n; i++) {
        pos[2 * i] += xdisp * (pos[2 * i + 1] - miny) / yrange;
        pos[2 * i + 1] += ydisp * (pos[2 * i + 1] - miny) / yrange;
    }
}

static void or_norms(int n, const double *f, double *ff, double *fmax2)
{
    int i;
    double s = 0.0, m = 0.0;
    for (i = 0; i < n; i++) {
        double a = f[2 * i], b = f[2 * i + 1];
        double q = fma(a, a, b * b);
        s += q;
        if (q > m) m = q;
    }
    *ff = s;
    *fmax2 = m;
}

/* run_fire_md, large_strain_relax.cu:596-762.  Same step order: drift and
 * half kick with the old force, new force, half kick, P = v.F, |v|, |F|,
 * mix v <- v - alpha v + (alpha |v|/|F|) F, adapt dt/alpha (or zero v when
 * P <= 0), then the convergence test on the force of this step.  The RMS
 * criterion is the reference's |F|/sqrt(2N) (:716); the max criterion
 * (extension) is the largest node force magnitude. */
typedef struct {
    const or_net *net;
    double ks, kb;
} or_net_field;

static void or_net_force(const void *ctx, const double *pos, double *force)
{
    const or_net_field *c = (const or_net_field *)ctx;
    or_forces(c->net, c->ks, c->kb, pos, force);
}

int or_fire(const or_net *net, double ks, double kb, double *pos,
            const double fp[5], int nmin, double mass, double fcut,
            int max_steps, int criterion, double report[4], int *steps)
{
    or_net_field c = {net, ks, kb};
    int rc = or_fire_core(net->n, or_net_force, &c, pos, fp, nmin, mass, fcut, max_steps,
                          criterion, &report[0], &report[3], steps);
    if (rc == OR_E_ALLOC) return rc;
    or_energies(net, ks, kb, pos, &report[1], &report[2]);
    return rc;
}

int or_fire_core(int n, or_force_fn fn, const void *ctx, double *pos, const double fp[5],
                 int nmin, double mass, double fcut, int max_steps, int criterion,
                 double *rms, double *fmax, int *steps)
{
    int ndof = 2 * n, i, since = 0, step = 0;
    double a0 = fp[0], fa = fp[1], dtmax = fp[2], finc = fp[3], fdec = fp[4];
    double alpha = a0, dt = dtmax, dtm = 0.5 * dt / mass, dtsq = 0.5 * dt * dt / mass;
    double sqrt_dof = sqrt((double)ndof), ff, fmax2, fmag;
    double *v = (double *)calloc((size_t)ndof, sizeof(double));
    double *f = (double *)malloc(sizeof(double) * (size_t)ndof);
    if (!v || !f) { free(v); free(f); return OR_E_ALLOC; }
    fn(ctx, pos, f);
    or_norms(n, f, &ff, &fmax2);
    fmag = sqrt(ff);
    while (step < max_steps) {
        double P = 0.0, vv = 0.0, vm, c;
        step++;
        for (i = 0; i < ndof; i++) {
            pos[i] = fma(dt, v[i], pos[i]);
            pos[i] = fma(dtsq, f[i], pos[i]);
            v[i] = fma(dtm, f[i], v[i]);
        }
        fn(ctx, pos, f);
        for (i = 0; i < ndof; i++) {
            v[i] = fma(dtm, f[i], v[i]);
        }
        for (i = 0; i < ndof; i++) {
            P = fma(v[i], f[i], P);
            vv = fma(v[i], v[i], vv);
        }
        or_norms(n, f, &ff, &fmax2);
        vm = sqrt(vv);
        fmag = sqrt(ff);
        c = alpha * vm / fmag;
        for (i = 0; i < ndof; i++) {
            v[i] = fma(-alpha, v[i], v[i]);
            v[i] = fma(c, f[i], v[i]);
        }
        if (P > 0) {
            since++;
            if (since > nmin) {
                dt = fmin(dt * finc, dtmax);
                dtm = 0.5 * dt / mass;
                dtsq = 0.5 * dt * dt / mass;
                alpha *= fa;
            }
        } else {
            since = 0;
            dt *= fdec;
            dtm = 0.5 * dt / mass;
            dtsq = 0.5 * dt * dt / mass;
            alpha = a0;
            memset(v, 0, sizeof(double) * (size_t)ndof);
        }
        if (criterion == OR_CRIT_MAX ? sqrt(fmax2) < fcut : fmag / sqrt_dof < fcut)
            break;
        if (!isfinite(ff)) break;
    }
    *rms = fmag / sqrt_dof;
    *fmax = sqrt(fmax2);
    *steps = step;
    free(v);
    free(f);
    return isfinite(ff) ? OR_OK : OR_E_ARG;
}

/* Nonlinear conjugate gradient (extension: the reference only names the
 * method, large_strain_types.hpp:31-34, "Fletcher-Reeve").  Fletcher-Reeves
 * directions, or Polak-Ribiere+ (beta = max(0, F.(F - F_old) / |F_old|^2), a
 * restart whenever that is not positive), restart to steepest descent when d
 * is not a descent direction,
 * secant line search on phi'(a) = -F(x + a d).d, trial points always formed
 * from the base point.  Each iteration first tests convergence. */
#define OR_CG_LS_MAX 12
#define OR_CG_ETA 1e-3
int or_cg(const or_net *net, double ks, double kb, double *pos, double fcut,
          int max_iters, int criterion, int pr, double report[4], int *iters)
{
    int ndof = 2 * net->n, i, it = 0, rc = OR_OK;
    double sqrt_dof = sqrt((double)ndof), ff, fmax2, gg, step_len = 1e-2;
    double *f = (double *)malloc(sizeof(double) * (size_t)ndof);
    double *d = (double *)malloc(sizeof(double) * (size_t)ndof);
    double *xt = (double *)malloc(sizeof(double) * (size_t)ndof);
    double *ft = (double *)malloc(sizeof(double) * (size_t)ndof);
    if (!f || !d || !xt || !ft) { rc = OR_E_ALLOC; goto done; }
    or_forces(net, ks, kb, pos, f);
    or_norms(net->n, f, &ff, &fmax2);
    gg = ff;
    memcpy(d, f, sizeof(double) * (size_t)ndof);
    while (it < max_iters) {
        double fd = 0.0, dd = 0.0, g0, a_prev, g_prev, a, g = 0.0, ggn, fof, beta;
        int ls;
        if (criterion == OR_CRIT_MAX ? sqrt(fmax2) < fcut : sqrt(ff) / sqrt_dof < fcut)
            break;
        if (!isfinite(ff)) { rc = OR_E_ARG; break; }
        it++;
        for (i = 0; i < ndof; i++) {
            fd = fma(f[i], d[i], fd);
            dd = fma(d[i], d[i], dd);
        }
        if (!(fd > 0)) {              /* not a descent direction: restart */
            memcpy(d, f, sizeof(double) * (size_t)ndof);
            fd = gg;
            dd = gg;
        }
        g0 = -fd;
        a_prev = 0.0;
        g_prev = g0;
        a = step_len / sqrt(dd);
        for (ls = 0; ls < OR_CG_LS_MAX; ls++) {
            double fdt = 0.0, an;
            for (i = 0; i < ndof; i++) xt[i] = fma(a, d[i], pos[i]);
            or_forces(net, ks, kb, xt, ft);
            for (i = 0; i < ndof; i++) fdt = fma(ft[i], d[i], fdt);
            g = -fdt;
            if (fabs(g) <= OR_CG_ETA * fabs(g0)) break;
            an = a - g * (a - a_prev) / (g - g_prev);
            if (!(an > 0) || !isfinite(an)) an = 2.0 * a;
            if (an > 4.0 * a) an = 4.0 * a;
            a_prev = a;
            g_prev = g;
            a = an;
        }
        if (ls == OR_CG_LS_MAX) {
            /* last trial point was not evaluated for the final a: evaluate */
            for (i = 0; i < ndof; i++) xt[i] = fma(a, d[i], pos[i]);
            or_forces(net, ks, kb, xt, ft);
        }
        step_len = a * sqrt(dd);
        /* Polak-Ribiere numerator's cross term F_new . F_old */
        fof = 0.0;
        for (i = 0; i < ndof; i++) fof = fma(ft[i], f[i], fof);
        memcpy(pos, xt, sizeof(double) * (size_t)ndof);
        memcpy(f, ft, sizeof(double) * (size_t)ndof);
        or_norms(net->n, f, &ff, &fmax2);
        ggn = ff;
        /* Fletcher-Reeves |F|^2/|F_old|^2 or Polak-Ribiere+ max(0, ...) */
        beta = pr ? fmax(0.0, (ggn - fof) / gg) : ggn / gg;
        for (i = 0; i < ndof; i++) d[i] = fma(beta, d[i], f[i]);
        gg = ggn;
    }
    report[0] = sqrt(ff) / sqrt_dof;
    report[3] = sqrt(fmax2);
    or_energies(net, ks, kb, pos, &report[1], &report[2]);
    *iters = it;
done:
    free(f); free(d); free(xt); free(ft);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Batch file (parse_batch_file :297-383, read_line :263-283).               */
/* Extensions ride in tokens the reference's sscanf ignores: line 2          */
/* "ks kb [kappa_lo kappa_hi seed]", line 4 "nmin mass fcut nmax [rms|max]    */
/* [fire|cg]".                                                                */
/* ------------------------------------------------------------------------ */
static int or_read_line(FILE *fp, or_line *ln, int *eof)
{
    if (*eof) return 0;
    while (!*eof) {
        or_getline(fp, ln, eof);
        if (ln->buf[0] == 0 || ln->buf[0] == '#') continue;
        break;
    }
    if (ln->buf[0] == 0 || ln->buf[0] == '#') return 0;
    return 1;
}

int or_parse_batch(const char *path, char *base, int base_cap, char *tag,
                   int tag_cap, or_batch *b)
{
    FILE *fp = fopen(path, "r");
    or_line ln = {0, 0};
    int eof = 0, nr, cnt = 0, ok = 0;
    char t1[4096], t2[4096], w1[64], w2[64];
    double mn, mx, inc, klo, khi, seed;
    char *save, *tok;
    memset(b, 0, sizeof(*b));
    base[0] = 0;
    tag[0] = 0;
    if (!fp) return OR_E_OPEN;
    if (!or_read_line(fp, &ln, &eof)) goto out;
    t1[0] = t2[0] = 0;
    nr = sscanf(ln.buf, "%4095s %4095s", t1, t2);
    if (nr < 1) goto out;
    snprintf(base, (size_t)base_cap, "%s", t1);
    if (nr > 1) snprintf(tag, (size_t)tag_cap, "%s", t2);
    if (!or_read_line(fp, &ln, &eof)) goto out;
    nr = sscanf(ln.buf, "%lf %lf %lf %lf %lf", &b->ks, &b->kb, &klo, &khi, &seed);
    if (nr < 2 || b->ks < 0 || b->kb < 0) goto out;
    if (nr == 5 && klo > 0 && khi >= klo) {
        b->kappa_lo = klo; b->kappa_hi = khi; b->kappa_seed = (uint64_t)seed;
    }
    if (!or_read_line(fp, &ln, &eof)) goto out;
    for (tok = strtok_r(ln.buf, " \t", &save); tok && cnt < 5;
         tok = strtok_r(NULL, " \t", &save)) {
        double x;
        if (sscanf(tok, "%lf", &x) == 1) b->fp[cnt++] = x;
    }
    if (cnt < 5) goto out;
    if (!or_read_line(fp, &ln, &eof)) goto out;
    w1[0] = w2[0] = 0;
    nr = sscanf(ln.buf, "%d %lf %lf %d %63s %63s", &b->nmin, &b->mass, &b->fcut,
                &b->nmax, w1, w2);
    if (nr < 4) goto out;
    if (strcmp(w1, "max") == 0 || strcmp(w2, "max") == 0) b->criterion = OR_CRIT_MAX;
    if (strcmp(w1, "cg") == 0 || strcmp(w2, "cg") == 0) b->method = 1;
    if (strcmp(w1, "cgpr") == 0 || strcmp(w2, "cgpr") == 0) b->method = 2;
    if (!or_read_line(fp, &ln, &eof)) goto out;
    nr = sscanf(ln.buf, "%lf %lf %lf", &mn, &mx, &inc);
    if (nr < 3 || mn < 0 || mx < 0 || inc < 0 || mx < mn) goto out;
    b->cmin = mn; b->cmax = mx; b->cinc = inc;
    if (!or_read_line(fp, &ln, &eof)) goto out;
    nr = sscanf(ln.buf, "%lf %lf %lf", &mn, &mx, &inc);
    if (nr < 3 || mn < 0 || mx < 0 || inc < 0 || mx < mn) goto out;
    b->smin = mn; b->smax = mx; b->sinc = inc;
    ok = 1;
out:
    fclose(fp);
    free(ln.buf);
    return ok ? OR_OK : OR_E_ARG;
}

/* get_dec_digits :188-198 and report_name :203-213 */
int or_dec_digits(int value)
{
    int d = 0;
    do { d++; value /= 10; } while (value > 0);
    return d;
}

void or_report_name(const char *base, int num_digits, int count,
                    const char *ext, char *out, int cap)
{
    int pad = num_digits - or_dec_digits(count), k, len;
    len = snprintf(out, (size_t)cap, "%s_", base);
    for (k = 0; k < pad && len < cap - 1; k++) out[len++] = '0';
    snprintf(out + len, (size_t)(cap - len), "%d.%s", count, ext);
}

static int or_relax(const or_net *net, const or_batch *b, double *pos,
                    double rep[4], int *steps)
{
    if (b->method == 1 || b->method == 2)
        return or_cg(net, b->ks, b->kb, pos, b->fcut, b->nmax, b->criterion, b->method == 2,
                     rep, steps);
    return or_fire(net, b->ks, b->kb, pos, b->fp, b->nmin, b->mass, b->fcut,
                   b->nmax, b->criterion, rep, steps);
}

static void or_log_line(FILE *lf, const double rep[4], int steps, double fcut,
                        int criterion)
{
    double measure = criterion == OR_CRIT_MAX ? rep[3] : rep[0];
    if (!lf) return;
    fprintf(lf, measure <= fcut ? "Convergence reached: " : "Convergence not reached: ");
    fprintf(lf, "Steps: %d, RMS Force: %12.8le\n", steps, rep[0]);
}

/* do_relaxation_run :792-1010: compression ramp, then shear ramp, each
 * increment an affine step of `inc` (the first one included) followed by a
 * relax from the previous state; log, edata and .pos outputs per step. */
int or_run_protocol(const char *net_file, const char *log_name,
                    const char *edata_name, const char *disp_base,
                    const or_batch *b, int strain_mode)
{
    or_net net;
    double *pos, range, smult, mn, mx, inc, cur = 0.0, max_comp, rep[4];
    int rc, steps, num_digits, disp_count = 1;
    FILE *lf, *ef;
    clock_t ct = clock();
    rc = or_read_network(net_file, &net);
    if (rc) return rc;
    if (b->kb > 0) {
        rc = or_get_triples(&net);
        if (rc) { or_free(&net); return rc; }
        if (b->kappa_hi > 0) or_set_fiber_kappa(&net, b->kappa_lo, b->kappa_hi, b->kappa_seed);
    }
    pos = (double *)malloc(sizeof(double) * 2 * (size_t)net.n);
    if (!pos) { or_free(&net); return OR_E_ALLOC; }
    memcpy(pos, net.pos, sizeof(double) * 2 * (size_t)net.n);
    range = net.height > 0 ? net.height : net.maxy - net.miny;
    smult = strain_mode ? range : 1.0;
    if ((b->cmax > 0 && b->cinc <= 0) || (b->smax > 0 && b->sinc <= 0) ||
        (net.height > 0 && b->cmax > 0)) {
        free(pos); or_free(&net); return OR_E_ARG;
    }
    num_digits = or_dec_digits(1 + (b->sinc > 0 ? (int)((b->smax - b->smin) / b->sinc) : 0));
    lf = fopen(log_name, "w");
    ef = fopen(edata_name, "w");
    mn = b->cmin * smult; mx = b->cmax * smult; inc = b->cinc * smult;
    if (mx > 0) {
        for (cur = mn; cur <= mx + OR_FLT_TOL; cur += inc) {
            or_affine(&net, pos, net.miny, range, 0.0, -inc);
            or_relax(&net, b, pos, rep, &steps);
            or_log_line(lf, rep, steps, b->fcut, b->criterion);
            if (ef) fprintf(ef, "%12.0lf\t%12.10lf\t%12.8le\t%12.8le\t%12.8le\n", 0.,
                            cur, rep[1], rep[2], rep[1] + rep[2]);
        }
    }
    max_comp = mx > mn ? cur - inc : mx;
    mn = b->smin * smult; mx = b->smax * smult; inc = b->sinc * smult;
    if (mx > 0) {
        for (cur = mn; cur <= mx + OR_FLT_TOL; cur += inc) {
            char name[4096];
            FILE *pf;
            int k;
            or_affine(&net, pos, net.miny, range, inc, 0.0);
            if (net.height > 0) net.le_shift += inc;
            or_relax(&net, b, pos, rep, &steps);
            or_log_line(lf, rep, steps, b->fcut, b->criterion);
            if (ef) fprintf(ef, "%12.10lf\t%12.10lf\t%12.10le\t%12.10le\t%12.10le\n",
                            cur, max_comp, rep[1], rep[2], rep[1] + rep[2]);
            or_report_name(disp_base, num_digits, disp_count++, "pos", name, sizeof(name));
            pf = fopen(name, "w");
            if (pf) {
                for (k = 0; k < net.n; k++)
                    fprintf(pf, "%2.10le\t%2.10le\n", pos[2 * k], pos[2 * k + 1]);
                fclose(pf);
            }
        }
    }
    ct = clock() - ct;
    if (lf) {
        fprintf(lf, "Elapsed time: %f seconds\n", ((float)ct) / CLOCKS_PER_SEC);
        fclose(lf);
    }
    if (ef) fclose(ef);
    free(pos);
    or_free(&net);
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Synthetic lattice.                                                        */
/* ------------------------------------------------------------------------ */
static uint64_t or_mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double or_uniform(uint64_t seed, uint64_t counter)
{
    uint64_t h = or_mix64(or_mix64(seed) + counter);
    return (double)(h >> 11) * 0x1.0p-53;
}

int or_lattice(int L, double p, uint64_t seed, int le_mode, or_net *net)
{
    int n = L * L, r, c, d, nb = 0, rc;
    double hgt = sqrt(3.0) * 0.5;
    double *pos;
    int *bond;
    signed char *wrap;
    if (L < 3 || (le_mode && (L & 1))) return OR_E_ARG;
    pos = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    bond = (int *)malloc(sizeof(int) * 6 * (size_t)n);
    wrap = (signed char *)malloc(6 * (size_t)n);
    if (!pos || !bond || !wrap) { free(pos); free(bond); free(wrap); return OR_E_ALLOC; }
    for (r = 0; r < L; r++)
        for (c = 0; c < L; c++) {
            pos[2 * (r * L + c)] = c + ((r & 1) ? 0.5 : 0.0);
            pos[2 * (r * L + c) + 1] = r * hgt;
        }
    for (r = 0; r < L; r++)
        for (c = 0; c < L; c++)
            for (d = 0; d < 3; d++) {
                int node = r * L + c, r2 = r, c2, wx = 0, wy = 0;
                if (d == 0) c2 = c + 1;
                else { r2 = r + 1; c2 = (d == 1) ? c + (r & 1) : c - 1 + (r & 1); }
                if (c2 == L) { c2 = 0; wx = 1; }
                if (c2 < 0) { c2 = L - 1; wx = -1; }
                if (r2 == L) {
                    if (!le_mode) continue;
                    r2 = 0;
                    wy = 1;
                }
                if (!(or_uniform(seed, (uint64_t)node * 3 + (uint64_t)d) < p)) continue;
                bond[2 * nb] = node;
                bond[2 * nb + 1] = r2 * L + c2;
                wrap[2 * nb] = (signed char)wx;
                wrap[2 * nb + 1] = (signed char)wy;
                nb++;
            }
    rc = or_net_from_arrays(n, nb, (double)L, le_mode ? L * hgt : 0.0, pos, bond, wrap, net);
    free(pos); free(bond); free(wrap);
    return rc;
}
