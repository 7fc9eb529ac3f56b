// nrelax_linear: drop-in for the reference program `sp_relax`
// (sparse_relax.cpp:1165-1234) in batch mode:
//   nrelax_linear -b <batch> [-g] [-i] [-n] [-s] [-z]
//   -b  batch file (4 lines: base [tag]; ks kb mu; compression min max inc;
//       shear min max inc)
//   -g  accepted: the solve always runs on the GPU
//   -i  itemize the energy file (stretch and bend energies apart)
//   -n  accepted: the solution is the minimum-norm one (no null-space part)
//   -s  loads are strains (compression x y extent, shear x x extent)
//   -z  displace the top row only instead of imposing the affine field
// The reference's interactive mode (no -b) is not offered.
#include <unistd.h>

#include <cstdio>

#include "../../include/nrelax.h"

int main(int argc, char **argv)
{
    const char *batch = nullptr;
    int strain_mode = 0, affine = 1, itemize = 0, c;
    while ((c = getopt(argc, argv, "b:ginsz")) != -1) {
        switch (c) {
        case 'b': batch = optarg; break;
        case 'i': itemize = 1; break;
        case 's': strain_mode = 1; break;
        case 'z': affine = 0; break;
        default: break;
        }
    }
    if (!batch) {
        std::fprintf(stderr, "No batch file was specified.\n");
        return 255;
    }
    return nr_linear_batch(batch, strain_mode, affine, itemize) == NR_OK ? 0 : 1;
}
