// generated by netgen.py -- do not edit
// the species and parameter names of network primordial2
const char* const kSpecies[] = {"H", "H+", "H-", "H2", "H2+", "H3+", "He", "He+", "He++", "HeH+", "e-"};
const char* const kParams[] = {"gamma_ad", "Tcmb"};
const char* const kElements[] = {"H", "He"};
const char* const kHeatNames[] = {""};
const char* const kCoolNames[] = {""};
const double kParamDefaults[] = {1.6666666666666667, 2.725};
const char* const kParamRequiredBy[] = {"", ""};
