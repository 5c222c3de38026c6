// care-gpu: command-line corrector on one MI355X (see include/care_host.h).
#include "../../include/care_host.h"
int main(int argc, char** argv) { return care_host_main(argc, argv); }
