// generated by netgen.py -- do not edit
// the species and parameter names of network xrays
const char* const kSpecies[] = {"H", "H+", "H-", "H2", "H2+", "He", "He+", "He++", "e-"};
const char* const kParams[] = {"J21", "gamma_ad", "Tcmb"};
const char* const kElements[] = {"H", "He"};
const char* const kHeatNames[] = {""};
const char* const kCoolNames[] = {""};
const double kParamDefaults[] = {0.0, 1.6666666666666667, 2.725};
const char* const kParamRequiredBy[] = {"", "", ""};
