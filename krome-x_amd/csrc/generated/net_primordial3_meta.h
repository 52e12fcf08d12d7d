// generated by netgen.py -- do not edit
// the species and parameter names of network primordial3
const char* const kSpecies[] = {"H", "H+", "H-", "H2", "H2+", "He", "He+", "He++", "e-"};
const char* const kParams[] = {"Tcmb", "tff", "opratio"};
const char* const kElements[] = {"H", "He"};
const char* const kHeatNames[] = {"CHEM", "COMPRESS"};
const char* const kCoolNames[] = {"H2", "CHEM", "CONT"};
const double kParamDefaults[] = {2.725, NAN, 3.0};
const char* const kParamRequiredBy[] = {"", "thermal.heating = ['COMPRESS']", ""};
