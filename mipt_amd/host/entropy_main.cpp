// entropy (translation-averaged S1/S2) on the batched HIP engine (drop-in for cpp/apps/entropy.cpp).
#include "mipt_hip.h"

int main(int argc, char **argv) { return mipt_hip_entropy_main(argc, argv); }
