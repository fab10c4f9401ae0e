/* zml's src/c.zig:1-4 imports "lapacke.h"; this library's LAPACKE subset is
 * declared in clapack.h. */
#ifndef ZMLHIP_LAPACKE_H
#define ZMLHIP_LAPACKE_H
#include "clapack.h"
#endif
