// generated by netgen.py -- do not edit
// the species and parameter names of network coolHCN
const char* const kSpecies[] = {"CN", "CO", "H", "H+", "H2", "H2O", "HCN", "HCO", "O", "OH", "e-"};
const char* const kParams[] = {"crate", "gamma_ad", "Tcmb"};
const char* const kElements[] = {"C", "H", "N", "O"};
const char* const kHeatNames[] = {""};
const char* const kCoolNames[] = {""};
const double kParamDefaults[] = {0.0, 1.6666666666666667, 2.725};
const char* const kParamRequiredBy[] = {"", "", ""};
