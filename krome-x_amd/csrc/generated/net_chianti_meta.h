// generated by netgen.py -- do not edit
// the species and parameter names of network chianti
const char* const kSpecies[] = {"C", "C+", "C++", "C+++", "C++++", "C+++++", "C++++++", "H", "H+", "e-"};
const char* const kParams[] = {"gamma_ad", "Tcmb"};
const char* const kElements[] = {"C", "H"};
const char* const kHeatNames[] = {""};
const char* const kCoolNames[] = {""};
const double kParamDefaults[] = {1.6666666666666667, 2.725};
const char* const kParamRequiredBy[] = {"", ""};
