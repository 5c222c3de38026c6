// Command-line front end of the CPU restatement (test infrastructure only).
#include "care_oracle.h"
int main(int argc, char** argv) { return orc_main(argc, argv); }
