/* Struct-layout table of the SW2 path's structures (TEST INFRASTRUCTURE: the checker of
 * include/adhk_ref.h).  Built twice: with -DADHK_LAYOUT_REFERENCE against the reference's own headers,
 * included in src/include/adh.h's order (oracle/Makefile target _ref/ref_layout), and against
 * include/adhk_ref.h (target ref_layout_adhk).  Prints "S <struct> <sizeof>" and
 * "F <struct>.<field> <offset> <size>" for the (struct, field) list of tests/golden/ref_layout_fields.h. */
#include <stddef.h>
#include <stdio.h>

#ifdef ADHK_LAYOUT_REFERENCE
#include <ctype.h>
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdbool.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
/* adh.h:35-106 without the external libraries, which these structures do not use */
#include "define.h"
#include "macro.h"
#include "model_codes.h"
#include "debug.h"
#include "header_tl_alloc.h"
#include "assert.h"
#include "sdt.h"
#include "scoverage.h"
#include "svect2d.h"
#include "svect.h"
#include "snode.h"
#include "stensor.h"
#include "selem_1d.h"
#include "selem_2d.h"
#include "selem_3d.h"
#include "squad.h"
#include "slist_items.h"
#include "smpi.h"
#include "smeteor.h"
#include "sflags.h"
#include "sstr_value.h"
#include "sgrid.h"
#include "sarray.h"
#include "tokens.h"
#include "sivar_position.h"
#include "smat_grid.h"
#include "smat_sw.h"
#include "smat_gw.h"
#include "smat_transport.h"
#include "smat_physics.h"
#include "slin_sys.h"
#include "dofmaps.h"
#include "sdvar.h"
#include "sfile.h"
#include "sseries.h"
#include "stestcase.h"
#include "smodel_super.h"
#include "sadh_def.h"
#include "smodel_design.h"
#else
#include "adhk_ref.h"
#endif

int main(void)
{
#define STRUCT(s) printf("S %s %zu\n", #s, sizeof(s));
#define FIELD(s, f) printf("F %s.%s %zu %zu\n", #s, #f, offsetof(s, f), sizeof(((s *)0)->f));
#include "ref_layout_fields.h"
    return 0;
}
