// generated by netgen.py -- do not edit
// the species and parameter names of network lowmetal
const char* const kSpecies[] = {"C", "C+", "H", "H+", "H-", "H2", "H2+", "He", "He+", "He++", "O", "O+", "Si", "Si+", "Si++", "e-"};
const char* const kParams[] = {"gamma_ad", "Tcmb"};
const char* const kElements[] = {"C", "H", "He", "O", "Si"};
const char* const kHeatNames[] = {""};
const char* const kCoolNames[] = {""};
const double kParamDefaults[] = {1.6666666666666667, 2.725};
const char* const kParamRequiredBy[] = {"", ""};
