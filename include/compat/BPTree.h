/* Compatibility header: the reference's BPTree.h call surface (BPTree.h:6-55),
 * implemented over the C ABI of libbptree_amd (include/bptree.h).
 *
 * Differences that callers can observe (all deliberate, see each function):
 *  - BTPage is opaque: `dev_globalPointer` is the tree handle, not a page pool.
 *  - Searches return their results (the reference returns 0 and frees them).
 *  - Grouped search returns results in the caller's order and leaves
 *    input_key unchanged; the pairing input_key[i] <-> result[i] holds as in
 *    the reference.
 *  - Range searches return a new[]'d array of inclusive counts |{k: s<=k<=e}|.
 *  - Failures print to stderr and return NULL instead of calling exit().
 */
#ifndef __INCLUDED_BPTree_H__
#define __INCLUDED_BPTree_H__

#include <vector>

#include "utils.h"

#define PAGE_ALLOCATION 200041 /* BPTree.h:6; no fixed pool here (exact-size allocation) */
#define PAGESIZE 5000          /* BPTree.h:8; the native leaf holds 8 (key, value) pairs */
#define GROUP_SEARCH_SIZE 33000
#define DOB_PAGE_SIZE 1
#define INVALID_DATA 0xffffffffffffffff

typedef struct bpt_tree BTPage; /* opaque: the tree handle */

/* BPTree.h:26-29. Builds from n sorted (key, data) pairs on the device and,
 * like the reference (BPTree.cu:210-211), frees both input buffers. The
 * groupList/groupPrefix leaf partition only sets page boundaries there and is
 * not needed by the flat layout. dev_btree_out receives a device ulli holding
 * the root id (0); dev_currentPage_in a device ulli holding the node count. */
void *genBPlusTree(ulli *devInputKey, ulli *devInput, ulli inputSize, std::vector<ulli> groupList,
                   std::vector<ulli> groupPrefix, ulli *&dev_btree_out, ulli &treeLevel_out,
                   ulli *&btreeIndex_out, ulli &bTreeSize_out, BTPage *&dev_globalPointer_in,
                   ulli *&dev_currentPage_in);

/* Releases what genBPlusTree allocated (no reference counterpart). */
void freeBPlusTree(BTPage *dev_globalPointer, ulli *dev_btree, ulli *dev_currentPage);

/* BPTree.h:36-37: host queries -> new[]'d host results (value or INVALID_DATA). */
void *SearchingBTreeGPUOnlyOneState(ulli *input_key, ulli inputSize, ulli *dev_Btree,
                                    BTPage *dev_globalPointer);

/* BPTree.h:40-41: same results, computed with on-device query grouping. */
void *SearchingBTreeGPUCombineStep(ulli *input_key, ulli inputSize, ulli *dev_Btree,
                                   BTPage *dev_globalPointer);

/* BPTree.h:47-55: host [start, end] pairs -> new[]'d host inclusive counts. */
void *SearchingBTreeGPURange(ulli *input_key_start, ulli *input_key_end, ulli inputSize,
                             ulli *dev_Btree, BTPage *dev_globalPointer);
void *SearchingBTreeGPURangeWithGrouping(ulli *input_key_start, ulli *input_key_end,
                                         ulli inputSize, ulli *dev_Btree, BTPage *dev_globalPointer);
void *SearchingBTreeGPURangeWithGroupingZorder(ulli *input_key_start, ulli *input_key_end,
                                               ulli inputSize, ulli *dev_Btree,
                                               BTPage *dev_globalPointer);

/* No reference counterpart: the range search with its results on the host.
 * The reference's state 3 (BtreeRangeSearch.cu:503-510, kernelGenerateSearchResult
 * :297-374) fills a device uni_result of inputSize * DOB_PAGE_SIZE slots, one
 * count per query, and frees it. Here the returned new[]'d array is that
 * uni_result layout (the inclusive count of query i in slot i * DOB_PAGE_SIZE),
 * and the ids come two-phase: output_offset (new[]'d, inputSize + 1 entries)
 * is the exclusive scan of the counts, and ids (new[]'d, output_offset[inputSize]
 * entries, at least one) holds the data of every key in [start, end], query
 * by query, in key order. groupMode: 0 plain, 1 2-D grouping, 2 Z-order. */
void *SearchingBTreeGPURangeResults(ulli *input_key_start, ulli *input_key_end, ulli inputSize,
                                    ulli *dev_Btree, BTPage *dev_globalPointer, int groupMode,
                                    ulli *&output_offset, ulli *&ids);
#endif
