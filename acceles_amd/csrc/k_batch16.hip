// k_batch16.hip -- batch_scores instantiations for blocks of 16 queries.
#include "k_batch_scores.hpp"

batch_fn get_batch_scores16(int E, bool is_signed, int bits) { return bscores_for<16>(E, is_signed, bits); }
