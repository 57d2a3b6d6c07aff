/* placeholder: filled in below */
#include "or_internal.h"
#include <stdlib.h>
void or_geom_clear_bodies(or_ctx *c) { (void)c; }
