/* eddy_oracle: command-line front end of the CPU oracle (test infrastructure),
 * same flags and outputs as the reference `eddy-gpu` binary. */
#include "eddy_oracle.h"
int main(int argc, char **argv) { return eo_main(argc, argv); }
