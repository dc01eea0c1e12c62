// mipt_probed, modes 0 and 1 (one, two or four references), on the batched HIP engine
// (drop-in for cpp/apps/mipt_probed.cpp).
#include "mipt_hip.h"

int main(int argc, char **argv) { return mipt_hip_probed_main(argc, argv); }
