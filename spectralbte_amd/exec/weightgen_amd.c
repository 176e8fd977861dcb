/* weightgen_amd -- drop-in for the reference's MPIWeightGenerator
 * (exec/MPIWeightGenerator.c):
 *
 *   weightgen_amd N glance beta lambda L_v
 *
 * run in a directory with Weights/.  glance 0 writes the Landau table on the
 * GPU.  All of the work is spbte_weightgen_main in libspectralbte_amd.so. */
#include "spbte_boltz.h"

int main(int argc, char **argv) { return spbte_weightgen_main(argc, argv); }
