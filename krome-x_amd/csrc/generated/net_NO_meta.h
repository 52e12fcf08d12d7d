// generated by netgen.py -- do not edit
// the species and parameter names of network NO
const char* const kSpecies[] = {"N", "N2", "NO", "O", "O2"};
const char* const kParams[] = {"gamma_ad", "Tcmb"};
const char* const kElements[] = {"N", "O"};
const char* const kHeatNames[] = {""};
const char* const kCoolNames[] = {""};
const double kParamDefaults[] = {1.6666666666666667, 2.725};
const char* const kParamRequiredBy[] = {"", ""};
