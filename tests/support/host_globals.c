/* Test host for libp2p_hip.so: defines the host-program state the drop-in
 * entry points read (the reference defines it in src/global.c).  Loaded
 * with RTLD_GLOBAL before libp2p_hip.so by tests/test_gpu_dropin.py. */
#include "../../include/p2p_hip.h"

Body *part;
Pack *leaf;
int first_leaf, last_leaf, NPART, NPACK, MAXNPACK, MAXLEAF;
int PROC_RANK, PROC_SIZE = 1, NEW_RANK;
P2PPack *packlist;
float SoftenScale, MASSPART, splitRadius, PIisq;
double INT2POS;
int domain_center_i[3];
float *acc[3];
int *task_t, *task_s;
