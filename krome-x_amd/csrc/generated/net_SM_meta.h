// generated by netgen.py -- do not edit
// the species and parameter names of network SM
const char* const kSpecies[] = {"H", "H2", "H2O", "O", "O2", "OH"};
const char* const kParams[] = {"gamma_ad", "Tcmb"};
const char* const kElements[] = {"H", "O"};
const char* const kHeatNames[] = {""};
const char* const kCoolNames[] = {""};
const double kParamDefaults[] = {1.6666666666666667, 2.725};
const char* const kParamRequiredBy[] = {"", ""};
