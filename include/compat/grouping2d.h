/* Compatibility header: the reference's grouping2d.h (grouping2d.h:6-10). */
#ifndef __INCLUDED_GROUPING_2D_H__
#define __INCLUDED_GROUPING_2D_H__

#include <vector>

#include "utils.h"

/* grouping2d.cu:24-31 */
int bucketIdentify2d(ulli input, ulli input2, int nPass);

/* grouping2d.cu:545-743: host inputs, new[]'d host outputs ordered by the
 * 2-D bucket of successive passes; group lists as the reference's. */
void genGrouping2d(ulli *input, ulli *input2, ulli inputSize, std::vector<ulli> &groupList,
                   ulli *&output, ulli *&output2, std::vector<ulli> &groupPrefix,
                   ulli Local_Group_Size);
#endif
