// nrelax: drop-in for the reference program `lsrelax`
// (large_strain_relax.cu:1047-1126): nrelax -b <batch> [-a] [-i] [-x bx] [-y by]
//   -b  batch file (6 lines, see parse_batch_file)
//   -j  an XML job file instead (the format large_strain_types.hpp declares:
//       several runs per network, FIRE or ConjGrad, report flags)
//   -a  absolute displacements instead of strains (the reference's -a falls
//       through into -b with no argument; here it does what it describes)
//   -i  accepted and ignored, as in the reference
//   -x/-y  the reference's block shape; here a hint for nodes per tile (x*y)
#include <unistd.h>

#include <cstdio>
#include <cstdlib>

#include "../../include/nrelax.h"

int main(int argc, char **argv)
{
    const char *batch = nullptr, *job = nullptr;
    int strain_mode = 1, bx = 0, by = 1, c;
    while ((c = getopt(argc, argv, "ab:ij:x:y:")) != -1) {
        switch (c) {
        case 'a': strain_mode = 0; break;
        case 'b': batch = optarg; break;
        case 'j': job = optarg; break;
        case 'i': break;
        case 'x':
            if (std::sscanf(optarg, "%d", &bx) < 1 || bx < 1) {
                std::fprintf(stderr, "Block dimensions must be positive.\n");
                bx = 0;
            }
            break;
        case 'y':
            if (std::sscanf(optarg, "%d", &by) < 1 || by < 1) {
                std::fprintf(stderr, "Block dimensions must be positive.\n");
                by = 1;
            }
            break;
        default: break;
        }
    }
    if (!batch && !job) {
        std::fprintf(stderr, "No batch file was specified.\n");
        return 255;
    }
    nr_batch_opts opts = {};
    opts.precision = NR_PREC_FP64;
    opts.tile_nodes = bx > 0 ? bx * by : 0;
    int rc = job ? nr_xml_batch(job, strain_mode, &opts) : nr_relaxation_batch(batch, strain_mode, &opts);
    return rc == NR_OK ? 0 : 1;
}
