// Per-level parameters file reader (include/params.h:18-111): rows of
// `level block_x block_t ntest sap_block_x sap_block_t` with global aggregate-block
// counts and SAP blocks per (virtual) rank.
#ifndef PARAMS_H
#define PARAMS_H
#include <string>

#include "variables.h"

void readParameters(const std::string& inputFile);

#endif
