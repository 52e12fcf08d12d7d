// generated by krome-x_amd/build.py: the networks of the multi
// library multi_c3 (C-interface index, name, Tgas evolved), in launch
// order
#include "net_cloud.h"
#include "net_primordialZ.h"
#include "net_primordial.h"
#define KX_MULTI_NAME "multi_c3"
#define KX_MULTI_COUNT 3
#define KX_MULTI_NETS(X) X(2, cloud, false) X(1, primordialZ, false) X(0, primordial, false)
#include "kx_multi.h"
