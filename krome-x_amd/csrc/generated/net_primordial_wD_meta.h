// generated by netgen.py -- do not edit
// the species and parameter names of network primordial_wD
const char* const kSpecies[] = {"D", "D+", "D-", "H", "H+", "H-", "H2", "H2+", "HD", "HD+", "He", "He+", "He++", "e-"};
const char* const kParams[] = {"gamma_ad", "Tcmb"};
const char* const kElements[] = {"D", "H", "He"};
const char* const kHeatNames[] = {""};
const char* const kCoolNames[] = {""};
const double kParamDefaults[] = {1.6666666666666667, 2.725};
const char* const kParamRequiredBy[] = {"", ""};
