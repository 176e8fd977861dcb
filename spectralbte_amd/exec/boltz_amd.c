/* boltz_amd -- drop-in for the reference's boltz_ executable (exec/boltz.c):
 *
 *   boltz_amd <input file> <output flags file>
 *
 * run in a directory holding input/ (the two files, mesh and species files),
 * Weights/, Data/ and, for restarts, Restart/, as boltz_ is.  All of the work
 * is spbte_boltz_main in libspectralbte_amd.so. */
#include "spbte_boltz.h"

int main(int argc, char **argv) { return spbte_boltz_main(argc, argv); }
