// Learned position model (placeholder until the model is attached).
#include "bpt_internal.hpp"

namespace bpt {

struct LearnedModel {};

void learned_free(LearnedModel *m) { delete m; }

uint64_t point_learned_ws(const bpt_tree *, uint64_t) { return 0; }

bpt_status point_learned(const bpt_tree *t, const void *, uint64_t, void *, void *, uint64_t, hipStream_t) {
    BPT_CHECK(t && t->learned, BPT_ERR_UNSUPPORTED, "no learned model attached to this tree");
    return BPT_ERR_UNSUPPORTED;
}

}  // namespace bpt
