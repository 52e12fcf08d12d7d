// generated by krome-x_amd/build.py: the networks of the multi
// library multi (C-interface index, name, Tgas evolved), in launch
// order
#include "net_primordial3.h"
#include "net_primordial.h"
#define KX_MULTI_NAME "multi"
#define KX_MULTI_COUNT 2
#define KX_MULTI_NETS(X) X(1, primordial3, true) X(0, primordial, false)
#include "kx_multi.h"
