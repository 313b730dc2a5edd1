/*
 * CPU-ORACLE TEST INFRASTRUCTURE — not part of the product.
 * Dynamic Smagorinsky restatement (subgrid.f90:261-414); filled in below.
 */
#include <stdio.h>
#include <stdlib.h>

#include "orc.h"

void orc_dsm(orc_state *s) {
    (void)s;
    fprintf(stderr, "orc_dsm: not restated yet\n");
    abort();
}
