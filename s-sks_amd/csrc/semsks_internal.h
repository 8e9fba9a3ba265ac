/* semsks_internal.h -- declarations shared by the library's translation units. */
#ifndef SEMSKS_INTERNAL_H
#define SEMSKS_INTERNAL_H
#include "semsks.h"
#ifdef __cplusplus
extern "C" {
#endif
/* records an error text for semsks_last_error() and returns st */
semsks_status semsks_set_error(semsks_status st, const char *fmt, ...);
#ifdef __cplusplus
}
#endif
#endif
