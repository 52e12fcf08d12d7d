// generated by netgen.py -- do not edit
// the species and parameter names of network primordialZ_UV
const char* const kSpecies[] = {"C", "C+", "CH", "CH2", "CH3", "CH4", "CO", "CO+", "CO2", "H", "H+", "H-", "H2", "H2+", "H2O", "H2O+", "H3+", "H3O+", "He", "He+", "He++", "HeH+", "O", "O+", "O2", "O2+", "OH", "OH+", "e-"};
const char* const kParams[] = {"redshift", "gamma_ad", "Tcmb"};
const char* const kElements[] = {"C", "H", "He", "O"};
const char* const kHeatNames[] = {""};
const char* const kCoolNames[] = {""};
const double kParamDefaults[] = {0.0, 1.6666666666666667, 2.725};
const char* const kParamRequiredBy[] = {"", "", ""};
