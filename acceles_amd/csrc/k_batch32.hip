// k_batch32.hip -- batch_scores instantiations for blocks of 32 queries.
#include "k_batch_scores.hpp"

batch_fn get_batch_scores32(int E, bool is_signed, int bits) { return bscores_for<32>(E, is_signed, bits); }
