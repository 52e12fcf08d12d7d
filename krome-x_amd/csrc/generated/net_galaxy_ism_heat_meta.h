// generated by netgen.py -- do not edit
// the species and parameter names of network galaxy_ism_heat
const char* const kSpecies[] = {"H", "H+", "H-", "H2", "H2+", "He", "He+", "He++", "e-"};
const char* const kParams[] = {"crate", "G0", "d2g", "gamma_ad", "Tcmb"};
const char* const kElements[] = {"H", "He"};
const char* const kHeatNames[] = {"CHEM", "PHOTO"};
const char* const kCoolNames[] = {"H2", "CHEM"};
const double kParamDefaults[] = {0.0, 0.0, 0.01, 1.6666666666666667, 2.725};
const char* const kParamRequiredBy[] = {"", "", "", "", ""};
