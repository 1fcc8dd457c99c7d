// Internal declarations shared by the host-side sources of the PatchTable library.
#ifndef PT_INTERNAL_H
#define PT_INTERNAL_H

#include "../../include/patchtable_amd.h"

#include <cstdint>

// Records `msg` as the calling thread's last error and returns `code`.
int pt_fail(int code, const char *msg);
int pt_failf(int code, const char *fmt, ...);

// Returns from the calling function when `expr` yields a status other than PT_OK.
#define TRY_PT(expr)                      \
    do {                                  \
        int rc_ = (expr);                 \
        if (rc_ != PT_OK) return rc_;     \
    } while (0)

// Padded-coordinate packing of the reference (XY_TO_INT / INT_TO_X / INT_TO_Y,
// patchtable.h:163-165).
#define PT_XY_TO_INT(x, y) ((x) | ((y) << 16))
#define PT_INT_TO_X(v) ((v) & 65535)
#define PT_INT_TO_Y(v) ((v) >> 16)

// Largest image sides the packing holds: x takes the low 16 bits (padded x <= 65535) and y the
// bits above them in an int32, so a padded y must stay below 2^15 or INT_TO_Y decodes it as
// negative. Heights (descriptor mode: descriptor height + 7) are therefore at most 32768.
#define PT_MAX_H 32768
#define PT_MAX_W (65536 - 8)
#define PT_SIZE_OK(h, w) ((h) <= PT_MAX_H && (w) <= PT_MAX_W)

// Empty cell marker (TABLE_UNUSED = 2^30 - 1, patchtable.h:109-112).
#define PT_TABLE_UNUSED ((1 << 30) - 1)

// Entries of the value -> bin map of each partition (PATCH_PARTITION_NMAP, patchtable.h:216).
#define PT_NMAP 16384

// Initial distance of an unvisited NNF entry (DIST_INFINITY, patchtable.h:173).
#define PT_DIST_INFINITY 100000

#endif
