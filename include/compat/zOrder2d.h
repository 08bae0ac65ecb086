/* Compatibility header: the reference's zOrder2d.h (zOrder2d.h:6-8). */
#ifndef __INCLUDED_ZORDER_2D_H__
#define __INCLUDED_ZORDER_2D_H__

#include "utils.h"

/* zOrder2d.cu:65-171: host inputs; new[]'d host Z-order keys. */
void genGrouping2dZ(ulli *input, ulli *input2, ulli *&output, ulli inputSize);
#endif
