// nrelax_mesh: drop-in for the reference's network + FEM-gel composite
// program (network_mesh_relax.cu:1424-1503):
//   nrelax_mesh -b <batch> [-a] [-i] [-x bx] [-y by]
//   -b  batch file (6 lines; line 2 "ks kb bmod [smod [dilation]]")
//   -a  absolute displacements instead of strains (the reference's -a falls
//       through into -b with no argument; here it does what it describes)
//   -i  accepted and ignored, as in the reference
//   -x/-y  the reference's block shape; accepted and ignored (the launch
//       shapes follow the composite's sizes)
#include <unistd.h>

#include <cstdio>

#include "../../include/nrelax.h"

int main(int argc, char **argv)
{
    const char *batch = nullptr;
    int strain_mode = 1, c, v;
    while ((c = getopt(argc, argv, "ab:ix:y:")) != -1) {
        switch (c) {
        case 'a': strain_mode = 0; break;
        case 'b': batch = optarg; break;
        case 'x':
        case 'y':
            if (std::sscanf(optarg, "%d", &v) < 1 || v < 1)
                std::fprintf(stderr, "Block dimensions must be positive.\n");
            break;
        default: break;
        }
    }
    if (!batch) {
        std::fprintf(stderr, "No batch file was specified.\n");
        return 255;
    }
    return nr_mesh_relaxation_batch(batch, strain_mode, 0) == NR_OK ? 0 : 1;
}
