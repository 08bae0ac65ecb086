/* Compatibility header: reference utils.h:4-9 types (sc25index/B-tree-index-searching). */
#ifndef __INCLUDED_UTILS_H__
#define __INCLUDED_UTILS_H__

typedef unsigned long long int ulli;

typedef struct ulli2s {
    ulli a;
    ulli b;
} ulli2;

#endif
