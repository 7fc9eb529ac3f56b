/* Helpers shared by the oracle's C files (test infrastructure only). */
#ifndef NR_ORACLE_INTERNAL_H
#define NR_ORACLE_INTERNAL_H

#include <stdio.h>

#include "oracle.h"

#define OR_FLT_TOL 1e-10

/* std::getline emulation: the line without '\n'; *eof set when end of file
 * was reached while reading. */
typedef struct {
    char *buf;
    size_t cap;
} or_line;
int or_getline(FILE *fp, or_line *ln, int *eof);

/* sqrt(dx*dx + dy*dy) of bond (i, j) with image (W*wx, H*wy) */
double or_rest_length(const double *pos, double W, double H, int i, int j, int wx, int wy);

/* FIRE on n nodes with an arbitrary force field (the loop of run_fire_md);
 * fn writes the clamped force of pos into force (2n). */
typedef void (*or_force_fn)(const void *ctx, const double *pos, double *force);
int or_fire_core(int n, or_force_fn fn, const void *ctx, double *pos, const double fp[5],
                 int nmin, double mass, double fcut, int max_steps, int criterion,
                 double *rms, double *fmax, int *steps);

#endif
