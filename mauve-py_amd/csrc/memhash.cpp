// memhash.cpp -- placeholder C-ABI for SML/match entry points
#include "mh_internal.hpp"
using namespace mh;
extern "C" {
int mh_sml_create(mh_context*, const char*, uint64_t, uint64_t, mh_sml**) { set_error("not built"); return MH_ERR_UNSUPPORTED; }
int mh_sml_create_device(mh_context*, const void*, uint64_t, uint64_t, mh_sml**) { set_error("not built"); return MH_ERR_UNSUPPORTED; }
void mh_sml_free(mh_sml*) {}
int mh_sml_get_header(const mh_sml*, mh_sml_header*) { return MH_ERR_UNSUPPORTED; }
int mh_sml_read(mh_context*, const mh_sml*, uint64_t, uint64_t, mh_bmer*, uint64_t*) { return MH_ERR_UNSUPPORTED; }
int mh_sml_get_seed_mers(mh_context*, const mh_sml*, const uint32_t*, uint64_t, uint64_t*) { return MH_ERR_UNSUPPORTED; }
int mh_sml_find_mer(mh_context*, const mh_sml*, uint64_t, uint64_t*, int*) { return MH_ERR_UNSUPPORTED; }
int mh_sml_get_sequence(mh_context*, const mh_sml*, uint32_t*, uint64_t) { return MH_ERR_UNSUPPORTED; }
int mh_find_matches_pairwise(mh_context*, const mh_sml* const*, int, mh_match*, uint64_t, uint64_t*, mh_match_stats*) { return MH_ERR_UNSUPPORTED; }
}
