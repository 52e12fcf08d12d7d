// generated by netgen.py -- do not edit
// the species and parameter names of network ionsH2
const char* const kSpecies[] = {"C", "C+", "C++", "C+++", "C++++", "C+++++", "C++++++", "H", "H+", "H-", "H2", "H2+", "He", "He+", "He++", "Mg", "Mg+", "Mg++", "Mg+++", "Mg++++", "Mg+++++", "Mg++++++", "N", "N+", "N++", "N+++", "N++++", "N+++++", "N++++++", "Ne", "Ne+", "Ne++", "Ne+++", "Ne++++", "Ne+++++", "Ne++++++", "O", "O+", "O++", "O+++", "O++++", "O+++++", "O++++++", "S", "S+", "S++", "S+++", "S++++", "S+++++", "S++++++", "Si", "Si+", "Si++", "Si+++", "Si++++", "Si+++++", "Si++++++", "e-"};
const char* const kParams[] = {"gamma_ad", "Tcmb"};
const char* const kElements[] = {"C", "H", "He", "Mg", "N", "Ne", "O", "S", "Si"};
const char* const kHeatNames[] = {""};
const char* const kCoolNames[] = {""};
const double kParamDefaults[] = {1.6666666666666667, 2.725};
const char* const kParamRequiredBy[] = {"", ""};
