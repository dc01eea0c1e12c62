// mipt (three-site RDM record) on the batched HIP engine (drop-in for cpp/apps/mipt.cpp).
#include "mipt_hip.h"

int main(int argc, char **argv) { return mipt_hip_mipt_main(argc, argv); }
