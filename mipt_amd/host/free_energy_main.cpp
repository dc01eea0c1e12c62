// free_energy (measurement-record free energy) on the batched HIP engine
// (drop-in for cpp/apps/free_energy.cpp).
#include "mipt_hip.h"

int main(int argc, char **argv) { return mipt_hip_free_energy_main(argc, argv); }
