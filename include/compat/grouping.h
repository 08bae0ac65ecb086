/* Compatibility header: the reference's grouping.h (grouping.h:6-28).
 * The partitions run on the device; group lists follow the reference's
 * pass loop and host bookkeeping (grouping.cu:776-884, 1288-1363,
 * 1746-1791). Within a group the order is a stable sort's (the
 * reference's is arbitrary). The caller's groupList / groupPrefix are
 * replaced. Not declared here: radixGroupGPU (its declaration,
 * grouping.h:6-7, does not match its definition) and the __global__
 * kernels of grouping.h:9-14. */
#ifndef __INCLUDED_GROUPING_H__
#define __INCLUDED_GROUPING_H__

#include <vector>

#include "utils.h"

/* grouping.cu:25-29 */
int bucketIdentify(ulli input, int nPass);

/* grouping.h:17-18: host input; returns a device pointer (hipMalloc) to the
 * grouped keys that the caller frees; fills the group lists. */
void *genGrouping(ulli *input, ulli inputSize, std::vector<ulli> &groupList,
                  std::vector<ulli> &groupPrefix, ulli Local_Group_Size);

/* grouping.h:20-28: host inputs; new[]'d host outputs. */
void genGroupingWithOnedata(ulli *input, ulli *input_data, ulli inputSize, ulli *&output,
                            ulli *&output_data, std::vector<ulli> &groupList,
                            std::vector<ulli> &groupPrefix, ulli Local_Group_Size);
void genGroupingWithTwodata(ulli *input, ulli *input_data, ulli *input_data2, ulli inputSize,
                            ulli *&output, ulli *&output_data, ulli *&output_data2,
                            std::vector<ulli> &groupList, std::vector<ulli> &groupPrefix,
                            ulli Local_Group_Size);
#endif
