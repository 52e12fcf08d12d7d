// generated by netgen.py -- do not edit
// the species and parameter names of network photobin
const char* const kSpecies[] = {"C", "C+", "H", "H+", "H2", "He", "He+", "e-"};
const char* const kParams[] = {"crate", "gamma_ad", "Tcmb"};
const char* const kElements[] = {"C", "H", "He"};
const char* const kHeatNames[] = {""};
const char* const kCoolNames[] = {""};
const double kParamDefaults[] = {0.0, 1.6666666666666667, 2.725};
const char* const kParamRequiredBy[] = {"", "", ""};
