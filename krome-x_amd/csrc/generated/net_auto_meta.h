// generated by netgen.py -- do not edit
// the species and parameter names of network auto
const char* const kSpecies[] = {"C", "C+", "e-"};
const char* const kParams[] = {"gamma_ad", "Tcmb"};
const char* const kElements[] = {"C"};
const char* const kHeatNames[] = {""};
const char* const kCoolNames[] = {""};
const double kParamDefaults[] = {1.6666666666666667, 2.725};
const char* const kParamRequiredBy[] = {"", ""};
