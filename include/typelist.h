/* typelist.h -- the reference's typelist.h is an empty include guard; the
 * TypeEntryList declarations live in reshape.h, as there. */
#ifndef TYPELIST_H
#define TYPELIST_H
#include "reshape.h"
#endif
