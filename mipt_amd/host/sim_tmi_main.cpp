// sim_tmi on the batched HIP engine (drop-in for cpp/apps/sim_tmi.cpp).
#include "mipt_hip.h"

int main(int argc, char **argv) { return mipt_hip_sim_tmi_main(argc, argv); }
